/*
 * CPU oracle for the WCNS pseudo-time-step hot path -- TEST INFRASTRUCTURE ONLY.
 *
 * A plain-C restatement of the reference Fortran (zhanghong24/WCNS, its src/ tree),
 * one function per reference routine, in the reference's operation order, so
 * that the GPU library can be checked against it on identical inputs.  Only
 * tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it,
 * and only as the checker / CPU baseline, never as the product path.
 *
 * Array conventions (identical to the reference's Fortran bounds):
 *   "G3" field arrays cover i,j,k in -2..n+3; element (i,j,k) lives at
 *   (i+2) + (ni+6)*((j+2) + (nj+6)*(k+2)).  Conservative q/dq are stored
 *   variable-fastest, 5 doubles per node, like q(1:5,i,j,k).
 *   Grid coordinates x,y,z cover 1..n: (i-1) + ni*((j-1) + nj*(k-1)).
 */
#ifndef WCNS_ORACLE_H
#define WCNS_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
    /* namelist values, param.dat groups inflow/step/flowtype/technic */
    double moo, reynolds, attack, sideslip, tref, twall;
    double cfl, timedt_rate, efix, csrv, gama, prl, prt;
    int nvis, nlamtur, nscheme, nflux, ntmst;
    /* namelist force (param.dat:34-44) */
    double sref, lfref, xref, yref, zref;
    int nwholefield;
    /* SA model (nlamtur = 1): step timedt_turb, turbulence group kmaxlim,
     * interplate group xk, xb (NND limiter of the convection term) */
    double timedt_turb, kmaxlim, xk, xb;
    double lref;          /* namelist force: grid length unit (PreProcess.f90:93, 994) */
} o_params;

typedef struct o_case o_case;

o_case *o_case_create(const o_params *p);
void    o_case_destroy(o_case *c);
int     o_block_add(o_case *c, int ni, int nj, int nk,
                    const double *x, const double *y, const double *z);
/* One region line of the .inp file (signed s_st/s_ed exactly as written),
 * plus for bctype<0 the target line (t_st/t_ed signed, nbt 1-based, ibcwin). */
int     o_region_add(o_case *c, int nb, const int s_raw[6], int bctype,
                     const int t_raw[6], int nbt, int ibcwin);
int     o_setup(o_case *c);           /* bc order, images, GCL metrics, freestream, groups */
int     o_set_prim(o_case *c, int nb, const double *r, const double *u, const double *v,
                   const double *w, const double *p);        /* G3 arrays */
int     o_get_field(const o_case *c, int nb, int which, double *out); /* G3 (or 5*G3 for q/dq) */
int     o_set_field(o_case *c, int nb, int which, const double *in);
long    o_g3_size(const o_case *c, int nb);
void    o_freestream(const o_case *c, double out[8]);  /* roo,uoo,voo,woo,poo,coo,too,visc */

/* one pseudo-time step in the order of Main.f90:85-109 */
int     o_boundary_all(o_case *c);
int     o_initial_q_dq_c_t(o_case *c);
int     o_timestep(o_case *c);
int     o_rhs_vis(o_case *c);
int     o_rhs_invis(o_case *c);
int     o_communicate_dq(o_case *c);
int     o_lhs(o_case *c);
int     o_communicate_pv(o_case *c);
int     o_residual(o_case *c, double *rms, double *resmax, int loc[5]);
int     o_step(o_case *c, int nsteps, double *rms_hist);
/* OpenMP threads of the pencil, node and hyperplane loops (default 1); the
 * results do not depend on it */
void    o_set_threads(o_case *c, int n);
long    o_counter(const o_case *c, int which);   /* 0 wcns fallbacks, 1 repairs */
/* Spalart-Allmaras one-equation model (turblence.f90), per block, as called
 * from r_h_s_vis before the mean-flow viscous flux (RHS.f90:22) */
int     o_turbulence(o_case *c, int nb);
/* calc_wall_dists_parallel (parallel.f90:4313-4468): nearest no-slip wall node */
int     o_wall_distance(o_case *c);
/* force + force_dif (PostProcess.f90:1415-1503, 1505-1680): cfx,cfy,cfz,cmx,cmy,cmz,cd,cl,xcp */
int     o_force(o_case *c, double out[9]);

/* image/jmage/kmage of region nr of block nb over its face box (s_st..s_ed) */
int     o_region_images(const o_case *c, int nb, int nr, int *img, int *jmg, int *kmg, int *box);
int     o_multipoint(const o_case *c, int *ngroups, int *npts, int *out, int cap);

/* pencil-level routines with the reference argument layout (for golden tests) */
void o_value_half_node(int n, int nmax, int ni, const double *q, double *q_half);
void o_flux_dxyz(int nl, int nmax, int ni, const double *f, double *df);
void o_flux_sw(const double *priml, const double *primr, double nx, double ny, double nz,
               double nt, double *f, double efix, double gama);
/* van Leer (flux.f90:395-501); -1 where gama p / rho < 0 */
int  o_flux_vanleer(const double *priml, const double *primr, double nx, double ny, double nz,
                    double nt, double *f, double efix, double gama);
int  o_flux_roe(const double *ql, const double *qr, double nx, double ny, double nz,
                double nt, double *f, double efix, double gama);
/* scheme 4 or 41; flux 1 (SW) or 3 (Roe); q is q(1:6,-2:nmax+3,1:2), trxyz(5,nmax), df(5,nmax) */
int  o_wcns_line(int scheme, int flux, int ni, int nmax, double efix, double gama,
                 const double *trxyz, const double *q, double *df);
int  o_wcns_line_ex(int scheme, int flux, int ni, int nmax, double efix, double gama,
                    const double *trx1, const double *q, const double *rpmin, double *df);
void o_value_half_node_45(int ni, const double *q, double *q_half);
void o_pave(const double *ql, const double *qr, double *qa, double gama);
void o_lxq(const double *prim, const double *gykb, const double *dq, double *f, double gama);
void o_rxq(const double *prim, const double *gykb, const double *dq, double *f, double gama);
void o_flux_roemx(const double *ql, const double *qr, double sx, double sy, double sz, double rpmin,
                  double *flr, double gama);
void o_mxdq_std(const double *prim, const double *gykb, const double *dq, double *f,
                double rad, int npn, double gama);

#ifdef __cplusplus
}
#endif
#endif
