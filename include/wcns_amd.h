/*
 * wcns_amd.h -- C-ABI of the MI355X (gfx950) WCNS pseudo-time-step library.
 *
 * The reference (zhanghong24/WCNS, Fortran 90 + MPI) runs its hot path as
 * argument-less subroutines over module globals.  This library exports one
 * entry point per such subroutine, operating on device-resident state held in
 * an opaque context.  All functions return 0 on success and a negative
 * WCNS_E_* code on failure; none of them aborts the process.
 *
 * Arrays crossing the boundary are plain host pointers:
 *   - grid coordinates x,y,z of a block: ni*nj*nk doubles, i fastest,
 *     node (i,j,k) (1-based) at (i-1) + ni*((j-1) + nj*(k-1))
 *     (PLOT3D order, reference src/PreProcess.f90:365-412);
 *   - ghosted fields ("G3", bounds -2..n+3 as in src/PreProcess.f90:720-786):
 *     (ni+6)*(nj+6)*(nk+6) doubles, node (i,j,k) at
 *     (i+2) + (ni+6)*((j+2) + (nj+6)*(k+2)).
 * Inside the library every field is a separate HBM array (structure of
 * arrays) in that same ghosted order.
 */
#ifndef WCNS_AMD_H
#define WCNS_AMD_H

#ifdef __cplusplus
extern "C" {
#endif

#define WCNS_OK               0
#define WCNS_E_ARG           -1   /* bad argument, index or size               */
#define WCNS_E_STATE         -2   /* call out of order (e.g. before setup)      */
#define WCNS_E_HIP           -3   /* a HIP runtime call failed                  */
#define WCNS_E_NOMEM         -4   /* device allocation failed                   */
#define WCNS_E_BC            -5   /* unsupported boundary type (BC.f90:201-205) */
#define WCNS_E_JACOBIAN      -6   /* non-positive cell volume (PreProcess.f90:2637-2666) */
#define WCNS_E_ROE           -7   /* Roe average c^2 <= 0 (flux.f90:179-182)     */
#define WCNS_E_REPAIR        -8   /* too many repaired nodes (Sub2.f90:781-785)  */
#define WCNS_E_UNSUPPORTED   -9   /* option of the reference not built here      */

/* Namelist values read by the hot path (reference src/param.dat groups inflow,
 * step, flowtype, turbulence, technic; src/PreProcess.f90:80-159). */
typedef struct {
    double moo, reynolds, attack, sideslip, tref, twall;
    double cfl, timedt_rate, efix, csrv, gama, prl, prt;
    int nvis;      /* 0 Euler, 1 Navier-Stokes                          */
    int nlamtur;   /* -1 laminar, 1 Spalart-Allmaras, 2 SST k-omega      */
    int nscheme;   /* 4 (WCNS_E_5), 41, 42, 45 or 47 (WCNS_E_5_4x)        */
    int nflux;     /* 1 Steger-Warming, 2 van Leer, 3 Roe, 4 RoeMx (with  */
                   /* nscheme 42 or 47, Convect.f90:34, 47)                */
    int ntmst;     /* 0 local time step                                  */
    /* namelist force (param.dat): reference area, length and moment centre */
    double sref, lfref, xref, yref, zref;
    int nwholefield; /* 1 whole body; 0 half model (forces doubled)      */
    /* namelist turbulence (param.dat, turblence.f90): SA time-step ratio,
     * mu_t/mu cap of max_control_turbulence, NND parameters of the SA
     * convection (Convect.f90:1281, turblence.f90:371-435) */
    double timedt_turb, kmaxlim, xk, xb;
    /* namelist force: length unit of the grid; the Reynolds number becomes
     * reynolds/lfref*lref (init_inflow, PreProcess.f90:994) */
    double lref;
} wcns_params;

/* Precision of the implicit operator: the residual, WCNS weights and the
 * conservative update are always FP64. */
#define WCNS_LHS_FP64 0
#define WCNS_LHS_FP32 1

typedef struct wcns_ctx wcns_ctx;

/* ---- context (no reference counterpart: the reference holds module globals,
 *      src/Globalvar.f90:284-423) ---------------------------------------- */
int wcns_create(wcns_ctx **ctx, int device, const wcns_params *prm);
int wcns_destroy(wcns_ctx *ctx);
const char *wcns_last_error(const wcns_ctx *ctx);

/* ---- setup: replaces read_plot3d_grid / read_bc_parallel / analyze_bc /
 *      set_grid_derivative (src/PreProcess.f90:365-412, src/parallel.f90:852-1301,
 *      src/PreProcess.f90:1618-1660) ------------------------------------- */
/* Adds a block; *nb receives its 0-based index. */
int wcns_add_block(wcns_ctx *ctx, int ni, int nj, int nk,
                   const double *x, const double *y, const double *z, int *nb);
/* One region line of the .inp file with the file's signs, s_raw = (s_st1,
 * s_ed1, s_st2, s_ed2, s_st3, s_ed3); for bctype -1 the target line t_raw,
 * target block nbt (1-based) and its region number ibcwin (1-based). */
int wcns_add_region(wcns_ctx *ctx, int nb, const int s_raw[6], int bctype,
                    const int t_raw[6], int nbt, int ibcwin);
/* GCL metrics on the device (GRID_DERIVATIVE_gcl, src/WCNSE5.f90:1478-1728),
 * boundary order (src/PreProcess.f90:2543-2573), interface maps
 * (src/parallel.f90:1185-1301) and multi-point groups (src/parallel.f90:4202-4311). */
int wcns_setup(wcns_ctx *ctx);
int wcns_set_lhs_precision(wcns_ctx *ctx, int prec);

/* ---- field exchange (host <-> device), G3 layout ------------------------ */
/* which: 0 r, 1 u, 2 v, 3 w, 4 p, 5 t, 7 dtdt, 8 sra, 9 srb, 10 src,
 *        11 srva, 12 srvb, 13 srvc, 14 visl, 16 vol, 17..25 kcx..ctz,
 *        100 dq (5 G3 arrays, variable-major), 102 rhs (the residual R);
 *        with a turbulence model (nlamtur = 1 SA, 2 SST) also 15 vist (mu_t),
 *        26 qke (mb_qke(:,:,:,1): SA nu~ or SST k), 27 dist (wall distance,
 *        nodes 1..n) and, for SST, 28 qkw (mb_qke(:,:,:,2), omega) */
int wcns_set_prim(wcns_ctx *ctx, int nb, const double *r, const double *u,
                  const double *v, const double *w, const double *p);
int wcns_get_field(wcns_ctx *ctx, int nb, int which, double *out);
int wcns_set_field(wcns_ctx *ctx, int nb, int which, const double *in);

/* ---- one pseudo-time step, stage by stage (src/Main.f90:85-109) --------- */
int wcns_boundary_all(wcns_ctx *ctx);        /* sub_tgh.f90:123-186     */
int wcns_initial_q_dq_c_t(wcns_ctx *ctx);    /* sub_tgh.f90:1-32        */
int wcns_timestep(wcns_ctx *ctx);            /* sub_tgh.f90:1270-1348   */
int wcns_r_h_s_vis(wcns_ctx *ctx);           /* RHS.f90:2-30            */
int wcns_r_h_s_invis(wcns_ctx *ctx);         /* RHS.f90:35-72           */
int wcns_communicate_dq(wcns_ctx *ctx);      /* parallel.f90:3020-3164  */
int wcns_l_h_s(wcns_ctx *ctx);               /* LHS.f90:1-34            */
int wcns_communicate_pv(wcns_ctx *ctx);      /* parallel.f90:3166-3326  */
/* residual_curve (PostProcess.f90:1195-1347): RMS and max of dq/(dtdt*vol)
 * over nodes 2..n-1 of the current dq; loc = i,j,k,m,nb (1-based).  Does not
 * touch the history below.  Synchronises the stream. */
int wcns_residual(wcns_ctx *ctx, double *rms, double *resmax, int loc[5]);

/* The same as wcns_residual, with its scratch (per-workgroup partials, the
 * per-block sums and the result) in a fixed device array compiled into the
 * library instead of buffers that wcns_setup allocates.  No counterpart in
 * the reference: it is for callers that cannot allocate device memory (code
 * under graph capture, tight serving loops, harnesses that forbid hipMalloc);
 * it calls no allocating routine.  Limit: at most 65536 monitor workgroups
 * (about 8 interior nodes per thread, 256 threads, at most 1024 per block) and
 * 4096 blocks; larger cases get WCNS_E_ARG.  Every context on one GPU shares
 * that GPU's array (1.7 MB), so calls that overlap on one GPU race; each call
 * rewrites every scratch word it reads, so nothing needs clearing between
 * calls.  Synchronises the stream. */
int wcns_residual_noalloc(wcns_ctx *ctx, double *rms, double *resmax, int loc[5]);

/* nsteps whole steps in the order of Main.f90:85-109, residual_curve
 * included: every step appends its monitor (rms, max, loc) to an on-device
 * history of the last 8192 steps, with no host synchronisation in between.
 * The host waits once per 8192 steps and at the end, where it checks the
 * device error counters; a Roe c^2 <= 0 or an update_limited abort
 * (WCNS_E_ROE / WCNS_E_REPAIR) is therefore reported at the end of that batch
 * of steps, not at the step where it occurred.  rms_hist (may be NULL)
 * receives each step's RMS. */
int wcns_step(wcns_ctx *ctx, int nsteps, double *rms_hist);

/* The same nsteps steps as wcns_step(ctx, nsteps, NULL), residual monitor
 * included, replayed from a HIP graph that holds two whole steps (captured on
 * first use, kept until the implicit-operator precision changes; an odd last
 * step runs as plain launches).  No counterpart in the reference.
 * Synchronises the stream at the end and checks the counters then. */
int wcns_step_graph(wcns_ctx *ctx, int nsteps);

/* The monitor history written by wcns_step / wcns_step_graph (the values
 * saveres2file writes to res.dat every nwerror steps, PostProcess.f90:1349-1378):
 * the last min(max, count, 8192) steps, oldest first; rms, resmax (max doubles
 * each) and loc (5*max ints: i,j,k,m,nb) may each be NULL; *count receives the
 * number of steps monitored since wcns_setup.  Synchronises the stream. */
int wcns_residual_history(wcns_ctx *ctx, int max, double *rms, double *resmax, int *loc, long long *count);

/* Tile-shape autotuner of the pipelined sweeps (no counterpart in the
 * reference): times the four sweeps of one step, reps times, for each
 * candidate (j, k) tile shape of at most 64 columns on the current state and
 * keeps the fastest for every block (the SA sweeps share it).  For the SGS
 * pair it also chooses between six helper wavefronts per tile and three that
 * form two operand parts each (two tiles per CU), unless the environment
 * variable WCNS_SGS_PAIR (0 or 1) fixed that at wcns_create.  Leaves the
 * primitive, SA and residual fields unchanged; the sweep unknowns dq are
 * scratch.  Returns the shape chosen and its time per sweep set in ms. */
int wcns_autotune(wcns_ctx *ctx, int reps, int *tj, int *tk, double *ms);

/* The pipelined sweeps' layout in use (no counterpart in the reference): the
 * first block's tile shape and whether the SGS helpers form two parts each. */
int wcns_sweep_layout(wcns_ctx *ctx, int *tj, int *tk, int *sgs_paired);

/* Force and moment coefficients over every no-slip wall (bctype 2 or 2x):
 * out = cfx, cfy, cfz, cmx, cmy, cmz, cd, cl, xcp.  Replaces force + force_dif
 * (src/PostProcess.f90:1415-1503, 1505-1680); the quad quadrature is getforce
 * (src/PostProcess.f90:1684-1796), the stress Sub1.f90:181-198 with the
 * one-sided wall derivatives of getuvwtder (Sub1.f90:383-480). */
int wcns_force(wcns_ctx *ctx, double out[9]);

/* wcns_force's coefficients appended to an on-device history of the last
 * 8192 records, with no host wait: what force() writes to force.dat every
 * nforce steps (PostProcess.f90:1415-1503).  No counterpart in the reference
 * (which writes each line at once). */
int wcns_force_record(wcns_ctx *ctx);
/* the last min(max, count, 8192) records, oldest first, 9 doubles each
 * (cfx, cfy, cfz, cmx, cmy, cmz, cd, cl, xcp); *count receives the number of
 * records since wcns_setup.  Synchronises the stream. */
int wcns_force_history(wcns_ctx *ctx, int max, double *out, long long *count);

/* ---- restart files in the reference's formats ---------------------------
 * wcns_save_flow writes `flowname` as savefile / write_flow_file_parallel do
 * (PostProcess.f90:1380-1396, parallel.f90:4073-4145): Fortran unformatted
 * sequential records, first (nout, wholetime, roo, uoo, voo, woo, poo, moo),
 * then per block r, u, v, w, p, t of every node i, j, k (i fastest); with a
 * turbulence model also flowname.turb as write_turbulence_parallel does
 * (parallel.f90:4676-4746): per block (qke(1..nlamtur), mu_t) per node.
 * wcns_load_flow reads both back as read_flow_file / read_turbulence_parallel
 * do (PreProcess.f90:1542-1616, parallel.f90:4603-4674): interior nodes only,
 * the ghosts keep their values (the next wcns_boundary_all sets them); T as
 * initial_t forms it from the primitives when t_from_file is 0 (what the
 * reference's restart does, PreProcess.f90:62), or as the file holds it when
 * t_from_file is 1: the temperature of the step before, which the wall
 * conditions read, so that a saved run resumes bit for bit.  nout and
 * wholetime (either may be NULL) receive the header values.  A file whose
 * records do not fit the case gives WCNS_E_ARG. */
int wcns_save_flow(wcns_ctx *ctx, const char *flowname, int nout, double wholetime);
int wcns_load_flow(wcns_ctx *ctx, const char *flowname, int t_from_file, int *nout, double *wholetime);

/* Tecplot output of post_processing's nplot choices (PostProcess.f90:37-55),
 * in the binary format (version 102) that the reference's mod_tecios writes
 * itself (parallel.f90:1-173), to the file `path` (the caller names it as
 * the reference does: tecname for the surfaces, its stem + "_wall.plt" /
 * "_vol.plt"):
 *   WCNS_TEC_SURFACE  tecout_bin_surface (PostProcess.f90:97-336): one zone
 *                     per boundary region of type 2, 3, 6 (not 4, 5, 7x or
 *                     interfaces); X Y Z R U V W P T M (+ vT1 (vT2) muT with
 *                     a turbulence model), the velocities of a wall taken
 *                     from the node inside it;
 *   WCNS_TEC_WALL     tecout_bin_wall (:338-582): one zone per wall region;
 *                     X Y Z cp, and with viscosity cfx cfy cfz cf qw (the
 *                     traction -2 s_lr (tau n) Re, its magnitude and the heat
 *                     flux, formed on the device).  The reference's qw reads a
 *                     local cp it never sets (:356, 475); here cp is the gas
 *                     model's 1/((gama-1) moo^2) (sub_tgh.f90:868);
 *   WCNS_TEC_VOLUME   tecout_bin_volume (:584-772): one zone per block. */
#define WCNS_TEC_SURFACE 0
#define WCNS_TEC_WALL    1
#define WCNS_TEC_VOLUME  2
int wcns_tecplot(wcns_ctx *ctx, const char *path, int kind);

/* device error/diagnostic counters: 0 WCNS positivity fallbacks
 * (WCNSE5.f90:2224-2245), 1 repaired nodes (Sub2.f90:661-767), 2 Roe c^2<=0 */
int wcns_get_counter(wcns_ctx *ctx, int which, long long *value);
int wcns_sync(wcns_ctx *ctx);

/* per-kernel timing with HIP events recorded on the context's own stream
 * (no counterpart in the reference, which only reads CPU_TIME, Main.f90:72).
 * wcns_timing(ctx, 1) clears and starts recording; wcns_timing_get returns
 * the summed elapsed time of a slot and the number of kernel launches in it. */
#define WCNS_T_BC           0
#define WCNS_T_INIT         1
#define WCNS_T_TIMESTEP     2
#define WCNS_T_VIS          3    /* node derivatives of u, v, w, T (viscous) */
#define WCNS_T_INV_I        4
#define WCNS_T_INV_J        5
#define WCNS_T_INV_K        6
#define WCNS_T_COMM_DQ      7
#define WCNS_T_SWEEP_LU_L   8
#define WCNS_T_SWEEP_LU_U   9
#define WCNS_T_SWEEP_GS_L  10
#define WCNS_T_SWEEP_GS_U  11
#define WCNS_T_UPDATE      12
#define WCNS_T_COMM_PV     13
#define WCNS_T_RESIDUAL    14
#define WCNS_T_SWEEP_STAGE 15   /* plane-major staging of the sweep operands */
#define WCNS_T_VIS_I       16   /* viscous flux derivative along i, j, k */
#define WCNS_T_VIS_J       17
#define WCNS_T_VIS_K       18
#define WCNS_T_SA_RHS      19   /* SA: gradients, sources, diffusion, NND convection, diagonal */
#define WCNS_T_SA_SWEEP_L  20   /* SA: scalar LU-SGS forward / backward sweep */
#define WCNS_T_SA_SWEEP_U  21
#define WCNS_T_SA_UPDATE   22   /* SA: update, mu_t, max_control */
#define WCNS_T_TRANSPOSE   23   /* i<->j transposed copies for the i-direction residuals */
#define WCNS_T_VIS_ROWS    24   /* viscous FLUX_DXYZ rows (WCNS_T_VIS_I/J/K: the half-node fluxes) */
#define WCNS_T_NSLOT       25
int wcns_timing(wcns_ctx *ctx, int on);
int wcns_timing_get(wcns_ctx *ctx, int slot, double *total_ms, long long *launches);

/* ---- tangent and adjoint of the residual (no counterpart in the reference) ----
 * R(P, X) of block nb is the right-hand side that wcns_initial_q_dq_c_t, the
 * mean-flow part of wcns_r_h_s_vis (mu_t held at its stored value: no
 * turbulence-model update) and wcns_r_h_s_invis form at the interior nodes from
 *   P  the primitives r, u, v, w, p of every node -2..n+3, ghosts included as
 *      wcns_boundary_all leaves them: 5 arrays of (ni+6)(nj+6)(nk+6) doubles,
 *      back to back, i fastest (the layout of wcns_get_field);
 *   X  the node coordinates x, y, z: 3 arrays of ni*nj*nk doubles (PLOT3D
 *      order), through the conservative metrics of GRID_DERIVATIVE_gcl
 *      (src/WCNSE5.f90:1478-1728).
 * R, dR and W are 5 arrays of ni*nj*nk doubles (interior nodes, i fastest).
 * Every array is DEVICE memory.  Each call leaves the block's primitives and
 * metrics set to (P, X) (the implicit operator's staged metrics follow; the
 * symmetry-plane normals and wall distances keep their set-up values) and the
 * residual of that state in place; calls are synchronous.  Blocks need more
 * than 4 nodes in every direction. */
int wcns_adj_residual(wcns_ctx *ctx, int nb, const double *P, const double *X, double *R);
/* tangent: dR = (dR/dP) dP + (dR/dX) dX (dP or dX may be NULL: zero); R may be NULL */
int wcns_adj_jvp(wcns_ctx *ctx, int nb, const double *P, const double *X, const double *dP,
                 const double *dX, double *R, double *dR);
/* adjoint: gP = (dR/dP)^T W (5 G3 arrays), gX = (dR/dX)^T W (3 arrays); either may be NULL */
int wcns_adj_vjp(wcns_ctx *ctx, int nb, const double *P, const double *X, const double *W,
                 double *gP, double *gX);
/* J = sum_c omega[c] * C[c], C the pressure part (src/PostProcess.f90:1560-1564)
 * of wcns_force's cfx, cfy, cfz, cmx, cmy, cmz summed over the block's no-slip
 * walls, and its gradient gP (dJ/dP, 5 G3 arrays), gX (dJ/dX); either may be NULL */
int wcns_adj_force(wcns_ctx *ctx, int nb, const double *P, const double *X, const double omega[6],
                   double *J, double *gP, double *gX);

/* ---- tangent and adjoint of a whole pseudo-time step (no counterpart in the reference) ----
 * G is one step of Main.f90:85-109 (FP64 implicit operator) as wcns_step
 * computes it: boundary_all, initial_q_dq_c_t, timestep_tgh, r_h_s_vis,
 * communicate_dq_npp, r_h_s_invis, communicate_dq_npp, l_h_s_tgh
 * (GS_PR_Single: lusgs_l/u and gs_pr_l/u, LHS.f90:381-577,1777-2046;
 * update_limited, Sub2.f90:603-796) and communicate_pv_npp.
 * The state is nstate arrays of the field pool (wcns_pool_layout: total
 * elements each, block nb's (ni+6)(nj+6)(nk+6) nodes from offsets[nb], i
 * fastest), DEVICE memory: r, u, v, w, p, then with nvis = 1 T (which the
 * no-slip wall reads one step late), then with a turbulence model nu~ or k,
 * mu_t, and for SST omega.  The derivatives are taken w.r.t. the first nvar
 * of them (r..p, and T when viscous) at the interior nodes: the turbulence
 * model's variables are carried along but held (frozen mu_t), ghost values
 * of P and dP are not read, and those of dPn and gP are zero; dP, dPn, Wn
 * and gP have nvar arrays, P and Pn nstate.  Every derivative is of the
 * arithmetic as computed, boundary conditions and interface averagings
 * included; a step in which a node is repaired is rejected
 * (WCNS_E_UNSUPPORTED).  Each call leaves the state at G(P). */
int wcns_pool_layout(wcns_ctx *ctx, long *total, long *offsets, int *nstate, int *nvar);
/* Pn = G(P), dPn = G'(P) dP (P NULL: the current state) */
int wcns_adj_step_jvp(wcns_ctx *ctx, const double *P, const double *dP, double *Pn, double *dPn);
/* gP = G'(P)^T Wn */
int wcns_adj_step_vjp(wcns_ctx *ctx, const double *P, const double *Wn, double *gP);
/* Reverse-mode sensitivity of J = sum_c omega[c] C[c] (the pressure part of
 * wcns_force's cfx..cmz over the no-slip walls) after nsteps steps from P0
 * (nstate arrays), with respect to P0's nvar arrays: checkpoints of the
 * whole state every `every` steps, each segment replayed from its
 * checkpoint and reversed.  Rejected (WCNS_E_ARG) when checkpoints and work
 * arrays need more than budget_gb (budget_gb <= 0: no limit); used_gb (may
 * be NULL) receives what they need. */
int wcns_adj_multistep(wcns_ctx *ctx, const double *P0, int nsteps, int every, const double omega[6],
                       double budget_gb, double *J, double *gP0, double *used_gb);

/* debugging/inspection of the host-side setup */
int wcns_region_images(wcns_ctx *ctx, int nb, int nr, int *img, int *jmg, int *kmg, int box[6]);
int wcns_multipoint(wcns_ctx *ctx, int *ngroups, int *npts, int *pts, int cap);

#ifdef __cplusplus
}
#endif
#endif
