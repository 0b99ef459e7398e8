/*
 * FP64 CPU oracle for the AMOEBA induced-dipole solve.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker that the HIP
 * product path is compared against; it is imported by tests/, by
 * __graft_entry__.smoke() and by bench.py's cpu_baseline leg, and by
 * nothing else.  The product (libthp) never links, loads or calls it.
 *
 * It restates, in plain C++17 + OpenMP and FP64 throughout, the Tinker-HP
 * 1.2 CPU path (reference tree v1.2/source):
 *   nblist / mlistcell ......... v1.2/source/nblist.f:17-247, image.f:113-140
 *   efld0_direct ............... v1.2/source/newinduce_pme.f:948-1389
 *   tmatxb_pme ................. v1.2/source/newinduce_pme.f:1392-1668
 *   efld0_recip ................ v1.2/source/newinduce_pme.f:1671-1917
 *   tmatxbrecip ................ v1.2/source/newinduce_pme.f:1919-2128
 *   bspline_fill_site/bsplgen .. v1.2/source/pmestuff.f:13-156
 *   grid_mpole_site ............ v1.2/source/pmestuff.f:163-274
 *   grid_uind_site ............. v1.2/source/pmestuff.f:280-400
 *   fphi_mpole_site ............ v1.2/source/pmestuff.f:453-631
 *   fphi_uind_site ............. v1.2/source/pmestuff.f:746-1037
 *   cmp_to_fmp/fphi_to_cphi .... v1.2/source/pmestuff.f:1088-1298
 *   moduli/bspline/dftmod ...... v1.2/source/kewald.f:672-812
 *   dampewald/dampthole(d) ..... v1.2/source/damping.f:17-215
 *   newinduce_pme2/inducepcg ... v1.2/source/newinduce_pme2.f:16-636
 *   epolar0c/epreal0c/eprecip .. v1.2/source/epolar.f:47-712
 *
 * Conventions at this C ABI (mirrors the reference's Fortran arrays in
 * C row-major order, 0-based atom indices):
 *   x,y,z[n]; rpole[n][13] = q, dx,dy,dz, Qxx,Qxy,Qxz,Qyx,Qyy,Qyz,Qzx,Qzy,Qzz
 *   (global frame, quadrupole already in the internal /3 convention);
 *   ef, mu, efi, dipfield: [n][2][3] with rhs 0 = d, 1 = p;  cphi: [n][10].
 *   Orthorhombic boxes only (box lengths bx,by,bz).
 */
#ifndef THP_ORACLE_H
#define THP_ORACLE_H
#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_sys orc_sys;

/* scaled pairs: ik[2*nscaled] (0-based, unordered), dpu[3*nscaled] = d,p,u scale */
orc_sys* orc_create(int n, const double* box, const double* x, const double* y,
                    const double* z, const double* rpole, const double* polarity,
                    const double* thole, int nscaled, const int* ik,
                    const double* dpu, double aewald, double cut, double buffer,
                    int nfft1, int nfft2, int nfft3, int order);
void orc_destroy(orc_sys* s);
void orc_set_coords(orc_sys* s, const double* x, const double* y, const double* z);
void orc_set_threads(int nthreads);

/* Verlet list (half list k>i, r^2 <= (cut+buffer)^2).  Returns pair count. */
long orc_nblist(orc_sys* s);
/* copy the pair set as sorted (i,k) rows, i<k; ik must hold 2*count ints */
void orc_nblist_pairs(orc_sys* s, int* ik);

void orc_efld0_direct(orc_sys* s, double* ef);
void orc_efld0_recip(orc_sys* s, double* cphi);
/* total permanent field: direct - grad(phi_recip) + self */
void orc_efld0_total(orc_sys* s, double* ef);
void orc_tmatxb(orc_sys* s, const double* mu, double* efi, int dodiag);
void orc_tmatxb_recip(orc_sys* s, const double* mu, double* dipfield);
/* full (alpha^-1 - T) mu including real, recip and self parts */
void orc_matvec_full(orc_sys* s, const double* mu, double* h);

/* PCG solve (inducepcg_pme2).  guess: 0 = alpha*ef (polgsf=0), 1 = mu0.
 * Outputs: mu[n][2][3], ep (kcal/mol), niter, resnrm.  Returns 0, or 1 when
 * politer is exhausted, 2 when p.h == 0 (the reference returns early). */
int orc_induce(orc_sys* s, double poleps, int politer, int guess,
               const double* mu0, double* mu, double* ef_out, double* ep,
               int* niter, double* resnrm, double* enehist);

/* polarization energy from the pairwise decomposition real + recip + self
 * (epolar0c / epreal0c / eprecip), for a given uind */
void orc_epolar_decomp(orc_sys* s, const double* uind, double* ereal,
                       double* erecip, double* eself);

/* building blocks exposed for the pinning tests */
void orc_bsplgen(double w, int order, double* theta /* [order][4] */);
void orc_bsmod(int nfft, int order, double* bsmod);
void orc_bsmod_raw(int nfft, int order, double* bsmod);
void orc_fft3d(int n1, int n2, int n3, double* data /* interleaved, x fastest */, int sign);
void orc_dampewald(double aewald, double r, double* dmpe /* [7] index 0..6 = orders 1..7 */);
void orc_damptholed(double pdi, double pdk, double pgamma, double r, int rorder, double* dmpik);
double orc_erfc(double x);
/* test hook: 0 = the reference's scaled-pair form; 1, 2 = planted mistakes */
void orc_plant_scaled_form(int form);

#ifdef __cplusplus
}
#endif
#endif
