/*
 * wcns_oracle_priv.h -- CPU oracle (TEST INFRASTRUCTURE ONLY): the view of one
 * block's arrays that the viscous and turbulence restatements work on.  Every
 * array except x, y, z (PLOT3D order, 1..n) and the 5-vectors has the G3 shape
 * (-2..n+3 in each direction, i fastest).
 */
#ifndef WCNS_ORACLE_PRIV_H
#define WCNS_ORACLE_PRIV_H
#include "wcns_oracle.h"

typedef struct {
    int ni, nj, nk;
    long g3;
    double *x, *y, *z;
    double *r, *u, *v, *w, *p, *t, *c;
    double *q, *dq, *rhs_nb;
    double *kcx, *kcy, *kcz, *etx, *ety, *etz, *ctx, *cty, *ctz, *vol;
    double *sra, *srb, *src, *srva, *srvb, *srvc, *dtdt, *visl, *vist;
    double *qke, *dist;
    double *qkw;          /* SST omega */
} oblock_view;

int o_view_block(o_case *c, int nb, oblock_view *v);
void o_view_params(const o_case *c, o_params *p, double *visc);
int o_view_threads(const o_case *c);
double o_view_muoo(const o_case *c);
void o_view_turb_inflow(const o_case *c, double *koo, double *omgoo);

static inline long OG(const oblock_view *b, int i, int j, int k)
{
    return (long)(i + 2) + (long)(b->ni + 6) * ((long)(j + 2) + (long)(b->nj + 6) * (long)(k + 2));
}
#endif
