"""MI355X-native WCNS pseudo-time-step hot path (WCNS-E5 residual + LU-SGS)."""
