/*
 * magic_hip.h — C ABI of libmagic_hip.so, the MI355X (gfx950) kernels behind
 * the pseudo-spectral dynamo time step.
 *
 * Plain C: device pointers are passed as `double*` / `int*` (complex arrays
 * are interleaved re/im doubles), sizes as ints, the HIP stream as `void*`
 * (a hipStream_t, NULL = default stream).  Every entry point is stream
 * ordered, allocates nothing, never synchronises and returns 0 on success,
 * -1 (MH_EINVAL) for invalid arguments and -2 (MH_ELAUNCH) for a launch
 * failure.  Callers own every buffer.
 *
 * The reference (rvilim/torch-magic, src/magic_torch/) has no native layer:
 * each entry point replaces torch-dispatched work behind a reference Python
 * function, cited per declaration.  Layouts:
 *   spectral fields   (L, N) complex, m-major lm order  (blocking.py:29-35)
 *   Fourier buffers   F[c][m][t][b] complex, t = interleaved N/S theta row
 *   grid              G[c][b][t][phi] real               (sht.py:298-348)
 */
#ifndef MAGIC_HIP_H
#define MAGIC_HIP_H

#ifdef __cplusplus
extern "C" {
#endif

#define MH_OK 0
#define MH_EINVAL -1
#define MH_ELAUNCH -2
#define MH_ETOOBIG -3 /* mh_gj_inverse_noalloc: sizes beyond its fixed work space */

/* ------------------------------------------------------------ Legendre */

/* P_lm, sin(theta) dP_lm/dtheta and their Gauss-weighted copies on the
 * northern nodes, written into the synthesis / analysis table layouts.
 * Replaces plms.build_plm_matrices (plms.py:158-186) and the padded/bucketed
 * SHT matrices of sht.py:84-289.  mp_info[(m*2+p)*4 + {0..3}] =
 * {K, syn_off, ana_off, K16} for parity p = (l-m)%2. */
int mh_plm_tables(int l_max, int n_m, int minc, int H, int Hp, const double* cos_t, const double* sin_t,
                  const double* wgt, const int* mp_info, double* synP, double* synD,
                  double* anaW, double* anaWd, void* stream);

typedef struct {
  const double* x1; /* (L, ldx) complex */
  const double* x2; /* mode 2 only */
  int mode;         /* 0: x1   1: l(l+1) x1   2: -x1 + or2[b] l(l+1) x2 */
  int pad;
} MhSynIn;

typedef struct {
  MhSynIn in[4][3]; /* [group][Q, S, T]; scalar synthesis uses in[g][0] */
  int ngroups, B, ldx, H, Hp;
  const int* mp_info;
  const int* lm_start;
  const double* synP;
  const double* synD;
  const double* or2;
  double* F;
  long c_stride, m_stride;
  int Bout;
  int prec;  /* 0: FP64 tables and MFMA; 1: FP32 tables (synP/synD hold floats), FP32 MFMA;
                2, 3: BF16 split tables of prec planes (synP/synD hold the bf16 tables of
                mh_tables_to_bf16), BF16 MFMA on the products of planes i, j with
                i + j < prec (3 or 6 products), FP32 accumulation */
  int fprec; /* 0: F holds doubles; 1: F holds floats (FP32 Fourier buffers, needs prec >= 1) */
  int waves; /* waves per workgroup (16 theta pairs each): 0 = default (4), 4 or 6 (prec <= 1) */
  int nct;   /* column tiles of 8 complex columns per workgroup: 0/1, or 2 (prec >= 1, vector) */
  int minc;  /* azimuthal symmetry: mode index m holds the order m * minc (0 reads as 1) */
  const int* mp16; /* prec >= 2: per (m, parity) {K32, synthesis offset, analysis offset, 0}
                      of the bf16 tables (elements) */
} MhSynArgs;

/* Legendre synthesis on FP64 MFMA: scalar (vector=0) replaces the bmm part of
 * scal_to_spat (sht.py:298-338); vector QST (vector=1) replaces the bmm part
 * of torpol_to_spat (sht.py:408-482), one launch for up to 4 field groups. */
int mh_leg_synth(int vector, const MhSynArgs* args, int n_m, void* stream);

typedef struct {
  const double* F;
  long c_stride, m_stride;
  int Bin, B, ncomp;
  int cf1[3], cf2[3];
  int pad0;
  double* outS[3];
  double* outT[3];
  int ldo, H, Hp, n_m; /* n_m: number of azimuthal orders */
  const int* mp_info;
  const int* lm_start;
  const double* anaW;
  const double* anaWd;
  const int* tiles; /* (m, p, k0) triples, k0 in steps of 96 degrees (one workgroup each) */
  int ntiles;
  int prec;  /* as MhSynArgs.prec (anaW/anaWd hold floats for 1, the bf16 tables for 2, 3) */
  int fprec; /* as MhSynArgs.fprec: F holds floats for 1 (needs prec >= 1) */
  int waves; /* waves per workgroup: 0 = default (6), 6 or 4 (4 only for prec >= 2); tiles
                must hold 16 * waves degrees */
  int minc;  /* as MhSynArgs.minc */
  const int* mp16; /* prec >= 2: as MhSynArgs.mp16 */
} MhAnaArgs;

/* Legendre analysis on FP64 MFMA: scalar replaces the bmm part of scal_to_SH
 * (sht.py:351-405); vector replaces spat_to_sphertor's products and the
 * division by l(l+1) (sht.py:575-667). */
int mh_leg_analysis(int vector, const MhAnaArgs* args, void* stream);

/* dst[i] = (float)src[i], i < n: the FP32 copies of the Legendre tables used by
 * the mixed-precision transforms (no reference counterpart: the reference's
 * transforms are FP64 only, sht.py:24-296). */
int mh_tables_to_f32(const double* src, float* dst, long n, void* stream);

/* The BF16 split tables of the split-precision transforms (no reference
 * counterpart): each FP64 table value x is written as np (2 or 3) bf16 planes
 * hi = bf16(x), then bf16 of each remainder, so that x = sum of the planes to
 * about 2^-(8 np) relative.  Layouts per (m, parity) block (offsets mp16[.. + 1],
 * mp16[.. + 2], in elements; planes consecutive within a block): synthesis
 * [plane][j < Hp][k < K32] (K32 = mp16[..]: degrees padded to 32, zero beyond
 * K), analysis [plane][k < K16][j < Hp] (zero for j >= H), so that a lane's 8
 * consecutive contraction indices of a BF16 MFMA operand are 16 contiguous
 * bytes.  Inputs: the FP64 tables of mh_plm_tables and their mp_info. */
int mh_tables_to_bf16(const double* synP, const double* synD, const double* anaW,
                      const double* anaWd, const int* mp_info, const int* mp16, int n_mp,
                      int H, int Hp, int np, unsigned short* synP16, unsigned short* synD16,
                      unsigned short* anaW16, unsigned short* anaWd16, void* stream);

/* ------------------------------------------------------------ FFT */
typedef struct {
  int Mh, nstage;
  int radix[16];
  int n_m, T, B, ncomp, RB, pad;
  const double* tw;  /* exp(-2 pi i k/Mh), complex, k < Mh */
  const double* tw2; /* exp(-2 pi i k/(2Mh)), complex, k < Mh */
  double* F;
  long Fc_stride, Fm_stride;
  int FB, pad0;
  double* G;
  long Gc_stride;
  const int* comp_scale; /* forward: nonzero -> multiply by ost2[t] */
  const double* ost2;
} MhFftArgs;

/* Azimuthal real FFT, norm="forward".  forward=0 replaces torch.fft.irfft in
 * sht.py:341,484-486; forward=1 replaces torch.fft.rfft (and the 1/sin^2
 * scaling) in sht.py:371,597-603. */
int mh_fft(int forward, const MhFftArgs* args, void* stream);

/* ------------------------------------------------------------ grid space */
typedef struct {
  const double* Gs;
  const double* Gv;
  long Gv_c_stride;
  double* Go;
  long Go_c_stride;
  int N, T, P, pad;
  const double* or2;
  const double* or4;
  const double* ost2;
  const double* valri2;
  const double* valhi2;
  double LFfac, cf2, af2;
  int l_mag, alf_damp;
  unsigned long long* cfl; /* 2N: per-level maxima (bit patterns of doubles) */
} MhNlArgs;

/* Nonlinear products and per-level Courant maxima: replaces get_nl
 * (get_nl.py:30-87) and the reductions of courant_check (courant.py:25-127). */
int mh_get_nl(const MhNlArgs* args, void* stream);

typedef struct {
  const double* Fs; /* s [1][M][T][N] complex (Fourier layout F[c][m][t][b]) */
  const double* Fv; /* [3][M][T][4N] complex, batch g*N + r, groups B, curl B, u, curl u */
  double* Fo;       /* [9][M][T][N-2] complex: Advr VSr VxBr Advt Advp VSt VSp VxBt VxBp */
  int N, T, M, Mh;
  int generic; /* 1: run the runtime-plan kernel even where a compiled plan exists */
  int fprec;   /* 0: Fs/Fv/Fo and the transforms in FP64; 1: FP32 (the pointers hold floats) */
  const double* tw;  /* exp(-2 pi i k/Mh) */
  const double* tw2; /* exp(-2 pi i k/(2Mh)) */
  const double *or2, *or4, *ost2, *valri2, *valhi2;
  double LFfac, cf2, af2;
  int l_mag, alf_damp;
  unsigned long long* cfl; /* 2N per-level maxima (bit patterns of doubles) */
  int rb;  /* radial levels per workgroup: 0 = the most the FFT plan fits, else 1..that */
  int pad;
} MhGridArgs;

/* Fused grid-space step, one workgroup per (theta row, RB radial levels):
 * inverse azimuthal FFTs of the 13 synthesised components in LDS, nonlinear
 * products and Courant maxima in registers, forward FFTs of the 9 products;
 * the grid never touches HBM.  Replaces irfft (sht.py:341,484-486), get_nl
 * (get_nl.py:30-87), courant_check (courant.py:25-127) and rfft with the
 * 1/sin^2 scaling (sht.py:371,597-603) of the radial loop. */
int mh_grid_step(const MhGridArgs* args, void* stream);

/* FFT plan of the grid step for half length Mh: out[0] = register radix,
 * out[1] = threads per level, out[2] = the most levels per workgroup (the
 * default level block; MhGridArgs.rb may ask for 1 .. out[2]).  Host only.
 * No counterpart in the reference: it exists for the launch-shape tuner. */
int mh_grid_plan(int Mh, int* out);

/* ------------------------------------------------------------ row operators */
typedef struct {
  const double* X;
  const int* rows;
  const double* MT;
  double* Y[6];
  int ldx, ldy, K, R, Rpad, rseg, nrows, pad;
} MhRowJob;

/* Y[lm][r] = sum_k M_g[r][k] X[lm][k] over a tile list of (job, row0, col0):
 * the per-degree implicit solves of all systems with the DCT-I and the radial
 * derivatives folded in (chunked_solve_complex algebra.py:271-310, costf
 * cosine_transform.py:18-45, get_dr/get_ddr/get_dddr radial_derivatives.py:
 * 18-68, the D123 matmul step_time.py:315-341).  wide = 0: tiles of 32 rows x
 * 64 columns (col0 multiples of 64); wide = 1: tiles of 32 rows x 128 columns
 * (col0 multiples of 128), fewer operand loads per MFMA.  Same results. */
int mh_rowop(const MhRowJob* jobs, const int* tiles, int ntiles, int wide, void* stream);

/* ------------------------------------------------------------ spectral side */
typedef struct {
  const int* l;
  const int* m;
  const double *dLh, *dTh2A, *dTh2S, *dTh3A, *dTh3S, *dTh4A, *dTh4S;
  int l_max, L;
} MhLmCoef;

typedef struct {
  const double *or1, *or2, *or3, *or4, *r2, *rgrav;
  int N, pad;
} MhRadCoef;

typedef struct {
  MhLmCoef lc;
  MhRadCoef rc;
  const double *Advr, *Advt, *Advp, *VSt, *VxBr, *VxBp;
  const double *dVSr_dr, *dVxBt_dr;
  const double *w, *dw, *z;
  double *ew, *ez, *ep, *es, *eb, *ej;
  double CorFac;
  int l_mag, pad;
} MhTdArgs;

/* Explicit terms: get_dwdt, get_dzdt, get_dpdt, get_dsdt, get_dbdt
 * (get_td.py:58-308) with finish_exp_entropy (update_s.py:191-210) and
 * finish_exp_mag (update_b.py:514-527). */
int mh_td(const MhTdArgs* args, void* stream);

typedef struct {
  const double *old, *impl, *expl0, *expl1;
  /* optional further terms of a BPR353 stage (weights wdev[8..11]); any term
   * pointer may be NULL (no term) */
  const double *t4, *t5, *t6, *t7;
} MhImexSys;

typedef struct {
  MhImexSys sys[6]; /* s, z, w, p, b, j */
  double *rs, *rz, *rb, *rj, *rwp;
  const int* m;
  int L, N;
  double wimp, wl1, we0, we1, bots_l0;
  int l_mag, pad;
  const double* wdev; /* optional: device weights {wimp, wl1, we0, we1, ...} replace the four above;
                         wdev[8..11] weigh the terms t4..t7 */
} MhRhsArgs;

/* IMEX right-hand sides with boundary rows and m=0 realness: replaces
 * _assemble_rhs_mega and the BC/WP-rhs assembly of _fused_lm_loop_fast
 * (step_time.py:186-219,239-280) and set_imex_rhs (time_scheme.py:75-90,
 * 240-262: CNAB2 and the BPR353 stages). */
int mh_rhs(const MhRhsArgs* args, void* stream);

typedef struct {
  double* rwp;
  const double* s;
  const double* ew0;
  const double* rgrav;
  const int* m;
  int L, N;
  double fac, BuoFac;
  int lm_l0, pad;
  const double* wdev; /* optional: device weights, wdev[4] replaces fac */
} MhBuoyArgs;

/* Implicit buoyancy coupling and the l=0 pressure right-hand side
 * (step_time.py:274-289). */
int mh_buoy(const MhBuoyArgs* args, void* stream);

typedef struct {
  MhLmCoef lc;
  MhRadCoef rc;
  const double *s, *ds, *d2s, *z, *d2z, *w, *dw, *ddw, *dddw, *p, *dp;
  const double *b, *ddb, *aj, *ddj;
  double *old_s, *impl_s, *old_z, *impl_z, *old_w, *impl_w, *old_p, *impl_p;
  double *old_b, *impl_b, *old_j, *impl_j;
  double opr, opm, BuoFac;
  int l_mag, lm_l0;
} MhImplArgs;

/* Stored old states and implicit terms (step_time.py:352-394,
 * update_b.py:619-631, setup_initial_state step_time.py:398-617). */
int mh_impl(const MhImplArgs* args, void* stream);

/* ------------------------------------------------------------ conducting inner core */
typedef struct {
  MhImexSys sys[2];      /* inner-core b, j: (L, n_ic) old, impl, expl0, expl1, t4..t7 */
  const double *rb, *rj; /* outer right-hand sides of mh_rhs, (L, N) */
  double *out_b, *out_j; /* coupled right-hand sides, (L, N + n_ic) */
  const int* m;          /* order of each mode (m = 0 rows are made real) */
  int L, N, n_ic, pad;
  double wimp, wl1, we0, we1;
  const double* wdev;    /* optional: device weights laid out as MhRhsArgs.wdev */
} MhIcRhsArgs;

/* Coupled outer/inner right-hand sides of the B and J systems of a conducting
 * inner core (update_b.py:655-668): 0, outer rows 1..N-2, 0 (ICB continuity),
 * 0 (ICB derivative match), inner IMEX right-hand side of levels 1..n_ic-1. */
int mh_ic_rhs(const MhIcRhsArgs* args, void* stream);

typedef struct {
  const double *b, *db, *ddb, *aj, *dj, *ddj; /* (L, n_ic) inner-core fields */
  double *old_b, *old_j;   /* (L, n_ic), or both NULL (no old state: BPR353 stages 1-3) */
  double *impl_b, *impl_j; /* (L, n_ic) */
  const double* dLh;       /* (L,) l(l+1) */
  const int* l;            /* (L,) degree */
  const double* O_r_ic;    /* (n_ic,) 1/r on the inner grid, 0 at the centre */
  int L, n_ic;
  double fac;              /* opm / sigma_ratio */
  double or2_icb;          /* 1/r_icb^2 */
} MhIcImplArgs;

/* Inner-core old states and implicit terms (get_mag_ic_rhs_imp,
 * update_b.py:803-873, and update_b.py:756-801): dLh or2(r_icb) f, and
 * fac dLh or2(r_icb) (f'' + 2(l+1) f'/r) on the bulk, (1 + 2(l+1)) f'' at the
 * centre; mode 0 and the ICB level untouched. */
int mh_ic_impl(const MhIcImplArgs* args, void* stream);

typedef struct {
  double* rot;         /* rotation state: [0] omega_ic, [1] old state, [2..5] implicit
                          terms, [6..9] explicit terms (Lorentz torque) of omega's equation */
  const double* wdev;  /* IMEX weights laid out as MhRhsArgs.wdev */
  int term[8];         /* rot indices of old, impl, expl0, expl1, t4, t5, t6, t7 in the
                          right-hand side (-1: absent), weighted as in MhRhsArgs */
  int expl_slot;       /* rot index the torque is written to */
  int impl_slot;       /* rot index the implicit term is written to */
  int store_old;       /* write rot[1] (CNAB2 always, BPR353 after the last stage) */
  int lm10;            /* row of the mode (l=1, m=0) */
  int L, N, n_ic, T, P;
  double c_z10_omega;  /* omega_ic = c_z10_omega z10(r_icb) */
  double c_dt_z10;     /* old state c_dt_z10 z10(r_icb) */
  double c_lorentz;    /* torque term c_lorentz * torque */
  double fac_lt;       /* LFfac 2 pi / n_phi */
  double or1_icb, or2_icb, or4_icb, r_icb;
  const double *br, *bp;      /* (T, P) ICB grid: radial, phi (sin(theta)-weighted) field */
  double *vt, *vp;            /* (T, P) output grid: rigid-rotation induction, zero */
  const double *gauss, *sin2; /* (T,) Gauss weight and sin^2(theta) of each grid row */
  const double *b_ic, *aj_ic; /* (L, n_ic) inner fields */
  double *e_b, *e_j;          /* (L, n_ic) inner explicit terms */
  const int* m;               /* (L,) order of each mode */
  const double* dLh;          /* (L,) l(l+1) */
  double* rhs_z;              /* (L, N) toroidal right-hand side */
  const double *z, *dz;       /* (L, N) toroidal velocity and its r-derivative */
} MhIcRotArgs;

/* Rotating conducting inner core (nRotIC = 1, no-slip ICB), reference
 * step_time.py:1126-1200, update_b.py:875-901, update_z.py:355-501,
 * step_time.py:597-617:
 *   mh_ic_torque: Lorentz torque fac_lt sum gauss br bp over the ICB grid, its
 *     explicit term c_lorentz torque into rot[expl_slot], and the rigid-rotation
 *     induction vt = or4 r_icb^2 sin^2(theta) omega br (vp = 0) on the grid;
 *   mh_ic_advect: inner explicit terms -omega or2(r_icb) i m l(l+1) f (0 at the ICB);
 *   mh_ic_omega_rhs: omega's IMEX right-hand side into rhs_z(l=1, m=0) at the ICB;
 *   mh_ic_omega_update: omega_ic = c_z10_omega z10, old state c_dt_z10 z10 and the
 *     implicit term -(2 z10 / r - dz10/dr) at the ICB from the solved z. */
int mh_ic_torque(const MhIcRotArgs* args, void* stream);
int mh_ic_advect(const MhIcRotArgs* args, void* stream);
int mh_ic_omega_rhs(const MhIcRotArgs* args, void* stream);
int mh_ic_omega_update(const MhIcRotArgs* args, void* stream);

/* ------------------------------------------------------------ anelastic step */
typedef struct {
  /* radial profiles of the reference state, (N,) each (radial_functions.py:71-167) */
  const double *beta, *dbeta, *rho0, *orho1, *otemp1, *dLtemp0;
  const double *visc, *dLvisc, *kappa, *dLkappa;
  const double* rgrav_rho; /* rho0 rgrav */
  double ViscHeatFac, OhmLossFac;
} MhAnelProf;

typedef struct {
  const double* Gs;     /* [b][t][phi]: s on levels 0..N-1, then l(l+1) z synthesised (cvr) */
  const double* Gv;     /* [c][g*N + b][t][phi]: g = 0 u, 1 du/dr, 2 (theta, phi) gradient
                           of the radial velocity; c = r, theta, phi as synthesised */
  long Gv_c_stride;
  const double* Gb;     /* [c][g*N + b][t][phi]: g = 0 B, 1 curl B (NULL: no field) */
  long Gb_c_stride;
  const double* Gd;     /* [c][b][t][phi]: azimuthal derivatives of u_theta, u_phi (raw) */
  long Gd_c_stride;
  double* Go;           /* [c][b-1][t][phi], bulk levels: 0 Advr 1 VSr 2 VxBr 3 heat 4 Advt
                           5 Advp 6 VSt 7 VSp 8 VxBt 9 VxBp */
  long Go_c_stride;
  int N, T, P, l_mag;
  const double *or1, *or2, *or4;
  const double *ost2, *cosn2; /* (T,) 1/sin^2 and cos/sin^2 of each grid row */
  const double *valri2, *valhi2;
  MhAnelProf prof;
  double LFfac, cf2, af2;
  int alf_damp, pad;
  unsigned long long* cfl; /* 2N: per-level Courant maxima (bit patterns of doubles) */
} MhNlAnelArgs;

/* Grid products of the anelastic step and per-level Courant maxima:
 * get_nl_anel (get_nl_anel.py:41-136: non-curl advection, Lorentz force,
 * v x B / rho0, v s, viscous and Ohmic heating) and courant_check
 * (courant.py:25-127, velocities weighted by 1/rho0^2 and Alfven speeds by
 * 1/rho0); the velocity and its derivatives are zero on the boundary levels
 * (step_time.py:1264-1268). */
int mh_nl_anel(const MhNlAnelArgs* args, void* stream);

typedef struct {
  const double* F;   /* Fourier buffer F[c][m][t][b] (complex as re, im) */
  long Fc_stride, Fm_stride;
  int FB;            /* columns of F */
  int c0;            /* first component read: out[k] = i m F[c0 + k] for k = 0, 1 */
  double* out;       /* F[k][m][t][b], k = 0, 1, with out's own strides */
  long Oc_stride, Om_stride;
  int OB;            /* columns of out */
  int B, n_m, T, minc, pad;
} MhDphiArgs;

/* Azimuthal derivative in Fourier space, i m F (m = mode index times minc), of
 * two components: with the inverse FFT it gives torpol_to_dphspat
 * (sht.py:725-743) before its 1/sin^2 factor. */
int mh_fourier_dphi(const MhDphiArgs* args, void* stream);

typedef struct {
  double* es;          /* (L, N) entropy explicit term of mh_td, updated in place */
  const double* heat;  /* (L, N) heating terms in spectral space */
  const double* orho1; /* (N,) */
  int L, N;
} MhEntropyAnelArgs;

/* es = (es + heat) / rho0: the heating source and finish_exp_entropy of the
 * anelastic step (step_time.py:1436-1440, update_s.py:191-210). */
int mh_entropy_anel(const MhEntropyAnelArgs* args, void* stream);

typedef struct {
  MhImplArgs base;   /* fields, old/impl outputs, opr, opm, BuoFac, l_mag, lm_l0 */
  MhAnelProf prof;
  const double* dz;  /* (L, N) dz/dr (the anelastic z implicit term has a first derivative) */
  int zero_edges;    /* zero the w, p implicit terms at l = 0 and the boundaries (set-up) */
  int pad;
} MhImplAnelArgs;

/* Old states and implicit terms of the anelastic step (update_s.py:265-268,
 * update_z.py:495-499, update_wp.py:500-531, step_time.py:398-617). */
int mh_impl_anel(const MhImplAnelArgs* args, void* stream);

typedef struct {
  double *z, *dz, *d2z;     /* (L, N) toroidal velocity and its derivatives, corrected */
  const double* r2q;        /* (N,) r^2 times the radial integration weights */
  const double *pz, *pdz, *pd2z; /* (N,) correction profiles rho0 r^2 and its derivatives */
  int l1m0, l1m1;           /* rows of (1, 0) and (1, 1) (-1: absent) */
  int N, do_z, do_e, pad;   /* axial / equatorial corrections */
  double cz;                /* 8/3 pi y10 / (c_moi_oc y10) */
  double ce;                /* 8/3 pi / c_moi_oc */
} MhAmArgs;

/* Angular-momentum corrections of z(l=1, m=0) (axial) and z(l=1, m=1)
 * (equatorial), update_z.py:301-331,450-475; one workgroup. */
int mh_am_correct(const MhAmArgs* args, void* stream);

typedef struct {
  double* rwp;        /* (L, 2N) (w, p) right-hand side; row lm_l0, entry N is written */
  const double* s;    /* (L, N) entropy */
  const double* u;    /* (N,) integral weights */
  int lm_l0, N;
} MhP0IntArgs;

/* The l=0 pressure's integral condition, rhs = sum u Re s(l=0)
 * (update_wp.py:431-436); one workgroup. */
int mh_p0_integral(const MhP0IntArgs* args, void* stream);

typedef struct {
  const double *w, *dw, *z, *b, *db, *aj;
  const int* lm_start;
  const double* or2;
  const double* qw;
  double* out;
  int l_max, N, by_m, l_mag;
  double LFfac;
  int minc; /* azimuthal symmetry: the orders are 0, minc, 2 minc, ... (0 reads as 1) */
  int pad;
  const double* orho1; /* (N,) 1/rho0 weighting the kinetic energies (anelastic), or NULL */
} MhEnergyArgs;

/* Kinetic / magnetic poloidal and toroidal energy per degree (by_m=0) or per
 * order (by_m=1), radially integrated: get_e_kin_full, get_e_mag_oc_full
 * (output.py:111-193) and the l/m aggregation of get_dlm (output.py:750-861);
 * out[q*(l_max+1) + i], q = kin pol, kin tor, mag pol, mag tor.  by_m=2: per
 * degree and split into bin = 2*(m == 0) + ((l + m) odd), out[(4q + bin)*(l_max+1)
 * + l] (16 rows), from which the axisymmetric and equatorially symmetric parts
 * of output.py:128-137,172-180 are sums.  by_m=3: radial profiles per degree,
 * not integrated (qw unused), out[(q*(l_max+1) + l)*N + r] for q = kin pol,
 * kin tor, kin pol over m > 0, kin tor over m > 0, then the same four magnetic
 * (8 planes): the e_lr, e_pol_lr and e_lr_c arrays of get_dlm
 * (output.py:768-800), halved (the reference's fac = 1/2). */
int mh_energy(const MhEnergyArgs* args, void* stream);

/* ------------------------------------------------------------ rebuild */
typedef struct {
  const double *R, *D1, *D2, *D3, *or1, *or2;
  double* out;
  int N, l_max, n_cheb, ktopv, kbotv, kind; /* kind 0 S 1 Z 2 B 3 J 4 WP 5 p0 6 all */
  double rnorm, wimp, opr, opm;
  const double* wimp_dev; /* optional: read the implicit weight here instead of wimp */
  const int* when;        /* optional: the kernel does nothing unless *when != 0 */
  double* out_wp;         /* kind 6: the (w,p) matrices (out then holds systems 0..nsys-1 */
  double* out_p0;         /*   stacked by system and degree); the l=0 pressure matrix */
  int nsys, pad;          /* kind 6: number of N x N systems (2: S, Z; 4: S, Z, B, J) */
} MhBuildArgs;

/* Per-degree implicit matrices: build_s_matrices (update_s.py:66-188),
 * build_z_matrices (update_z.py:89-240), build_wp_matrices
 * (update_wp.py:197-368), build_p0_matrix (update_wp.py:80-170),
 * _build_b_matrices_insulating (update_b.py:98-298). */
int mh_build_matrices(const MhBuildArgs* args, void* stream);

/* Batched equilibrated Gauss-Jordan inverse (precond 0 none, 1 rows, 2 rows
 * and columns), out = Fc inv(Fr A Fc) Fr; replaces the torch.linalg.inv calls
 * of the matrix builders.  work: nbatch*2n*n doubles, scal: nbatch*2n.
 * when: optional, as in MhBuildArgs. */
int mh_gj_inverse(const double* A, double* out, double* work, double* scal, int n, int nbatch,
                  int precond, int* info, const int* when, void* stream);

/* Capacity, in doubles, of the device work space of mh_gj_inverse_noalloc
 * (32 MiB, held on every device the library runs on until the process ends). */
#define MH_GJ_SCRATCH_DOUBLES (4L << 20)
#define MH_MAX_DEVICES 64

/* mh_gj_inverse without a work-space argument and without any allocation, for
 * callers that cannot allocate device memory (code under HIP graph capture,
 * serving loops that forbid hipMalloc).  No counterpart in the reference.
 * Systems of n > 208 use a fixed device array of MH_GJ_SCRATCH_DOUBLES doubles:
 * nbatch * (2n^2 + 2n) must fit, else MH_ETOOBIG and nothing is launched
 * (n <= 208 needs no work space, any nbatch).  All calls on one device share
 * that array: calls whose launches overlap on one device (several streams)
 * race; stream-ordered calls do not.  Results are bit-identical to
 * mh_gj_inverse's. */
int mh_gj_inverse_noalloc(const double* A, double* out, int n, int nbatch, int precond,
                          int* info, const int* when, void* stream);

typedef struct {
  const double* A;
  const double* B;
  double* C;
  long sAr, sAc, sAb, sBr, sBc, sBb, sCr, sCc, sCb;
  int M, Nn, K, batch;
  double alpha;
  const int* when; /* optional: the kernel does nothing unless *when != 0 */
} MhGemmArgs;

/* Batched strided FP64 GEMM C = alpha A B on MFMA. */
int mh_gemm_f64(const MhGemmArgs* args, void* stream);

/* ------------------------------------------------------------ step control */

/* Device-resident state of the time stepping of one run (read and written by
 * mh_step_control). */
typedef struct {
  double dt;        /* dt of the last step (the dt the next decision starts from) */
  double dtMax;
  double dt1, dt2;  /* dt history: last step, the step before */
  double wimp_lin0; /* implicit weight the operators were built for (NaN: none yet) */
  double time;
  int n_step;       /* steps taken */
  int rebuild;      /* 1 while the current step rebuilds the operators */
  int n_rebuild;
  int new_dt;       /* 1 when the current step's Courant decision changed dt */
  int sing_step;    /* first step whose rebuild met a zero pivot (0: none yet) */
  int sing_sys;     /* its system: row of the pivot-flag array */
  int sing_l;       /* its degree */
  int sing_info;    /* its first zero pivot (1-based) */
} MhStepState;

typedef struct {
  const unsigned long long* cfl; /* 2N per-level Courant maxima (FP64 bits): radial, then horizontal */
  const double* delxr2;          /* N: squared radial grid spacing (courant.py) */
  const double* delxh2;          /* N: squared horizontal grid spacing */
  MhStepState* state;
  double* w;                     /* out: wimp, wimp_lin[1], wexp[0], wexp[1], wimp_lin[0]*BuoFac */
  const double* dt_schedule;     /* optional: n_schedule step sizes replacing the decision */
  double* dt_log;                /* optional: ring of log_cap entries, dt of step n at (n-1) % log_cap */
  const int* info;               /* optional: pivot flags [info_rows][info_cols] of the last rebuild */
  int N, n_schedule, log_cap, info_rows, info_cols, pad;
  double alpha, BuoFac;
  /* scheme 1 (BPR353, time_scheme.py:157-326): the stage weights go to w_stage
   * (4 x 16 doubles: per stage [1, bimp_0, bexp_0, bexp_1, wimp_lin0 BuoFac, -, -, -,
   * bimp_1, bexp_2, bimp_2, bimp_3], all scaled by dt), wimp_lin0 = dt / 2 and
   * both dt history entries are dt; scheme 0 (CNAB2) writes w */
  double* w_stage;
  int scheme, pad2;
} MhStepCtlArgs;

/* The host decisions of one CNAB2 step taken on the device, so that a whole
 * step is one stream of kernels (one HIP graph): Courant limits from the
 * per-level maxima (courant.py:80-127), the hysteresis dt decision of
 * dt_courant (courant.py:130-155), the rebuild decision and dt roll of
 * one_step (step_time.py:1574-1637) and the CNAB2 weights of set_weights
 * (time_scheme.py:56-109).  state->rebuild and state->wimp_lin0 are the
 * `when` flag and `wimp_dev` of the rebuild launches that follow.  When the
 * previous step rebuilt, its pivot flags are scanned first and the first
 * singular system is kept in state->sing_* (the reference raises there,
 * update_s.py:152; the host raises when it reads the state). */
int mh_step_control(const MhStepCtlArgs* args, void* stream);

/* p[k*stride + i] = v for k < count, i < n, unless `when` (optional) points at 0. */
int mh_fill_f64(double* p, long n, int count, long stride, double v, const int* when,
                void* stream);

/* ------------------------------------------------------------ banded */

/* LINPACK dgbfa order pivoted band LU of a batch of (2kl+ku+1, n) band
 * matrices in place (negated multipliers, 1-based pivots, ZERO_TOLERANCE
 * 1e-15): prepare_band (algebra.py:371-457).  info[b] = 0 or first zero
 * pivot (1-based). */
int mh_band_factor(double* abd, int* pivot, int* info, int n, int kl, int ku, int nbatch,
                   void* stream);

/* dgbsl forward/back substitution for complex right-hand sides grouped by
 * degree, with row (and column) preconditioning: solve_band_real
 * (algebra.py:460-506) inside banded_solve_by_l (algebra.py:672-723).
 * rhs/sol: (L, n) complex rows (sol may alias rhs); rows of degree l use factor l.
 * rows: optional processing order of the L rows (grouped by degree, so that
 * the lanes of a wave share factors); kl, ku <= 4. */
int mh_band_solve(const double* abd, const int* pivot, const double* fac_row,
                  const double* fac_col, const int* lm2l, const int* rows, const double* rhs,
                  double* sol, int L, int n, int kl, int ku, void* stream);

/* ------------------------------------------------------------ double-curl (FD) step */
typedef struct {
  MhLmCoef lc;
  MhRadCoef rc;
  /* reference-state profiles, (N,) each; NULL selects the Boussinesq value
   * (beta, its derivatives and those of ln visc 0; orho1, rho0, visc 1) */
  const double *beta, *dbeta, *ddbeta, *orho1, *rho0, *visc, *dLvisc, *ddLvisc;
  const double* grav;      /* gravity profile rgrav (radial_functions.py) */
  /* mh_td_dc */
  const double *Advr;      /* (L, N) radial advection */
  const double *dAdvt_dr;  /* (L, N) d/dr (orho1 r^2 Advt), Advt zeroed on the boundaries */
  const double *w, *dw, *ddw, *dddw, *ddddw, *z, *dz, *s;
  double *ew, *ep;         /* explicit terms of w (written) and p (zeroed) */
  /* mh_buoy_dc */
  double* rwp;             /* (L, 2N) right-hand side: w half, then the pressure half */
  const double* ew0;       /* explicit w term of the step (l = 0 pressure) */
  const double* wdev;      /* optional: wdev[4] = wimp_lin[0] BuoFac replaces fac */
  /* mh_impl_dc (old_w, old_p, impl_p may be NULL) */
  double *old_w, *impl_w, *old_p, *impl_p;
  double CorFac, BuoFac, fac;
  int lm_l0, pad;
} MhDcArgs;

/* Double-curl poloidal equation (params.py:143, radial_scheme = "FD"):
 *   mh_td_dc: explicit w term get_dwdt_double_curl (get_td.py:91-147) with
 *     finish_exp_pol (update_wp.py:371-384), l = 0 in the standard form; p term 0;
 *   mh_buoy_dc: buoyancy wimp_lin[0] BuoFac dLh/r^2 rgrav s on rows 2..N-3 and
 *     the l = 0 pressure right-hand side (update_w_doublecurl.py:413-459);
 *   mh_impl_dc: get_w_old and get_w_impl (update_w_doublecurl.py:309-393), the
 *     pressure's old state and implicit terms zeroed. */
int mh_td_dc(const MhDcArgs* args, void* stream);
int mh_buoy_dc(const MhDcArgs* args, void* stream);
int mh_impl_dc(const MhDcArgs* args, void* stream);

/* out[i] = A0[i] + w A1[i], i < n; w = *wdev when wdev is not NULL; nothing
 * unless *when != 0 when `when` is not NULL.  Operators affine in the
 * implicit weight (the finite-difference systems, assembled once on the host)
 * take their new matrices from it on a dt change (build_s/z/w/b_matrices). */
int mh_affine_f64(const double* A0, const double* A1, double* out, long n, double w,
                  const double* wdev, const int* when, void* stream);

/* mh_band_factor / mh_band_solve that do nothing unless *when != 0 (`when`
 * may be NULL: always run); the finite-difference rebuild inside a captured
 * step uses them. */
int mh_band_factor_gated(double* abd, int* pivot, int* info, int n, int kl, int ku, int nbatch,
                         const int* when, void* stream);
int mh_band_solve_gated(const double* abd, const int* pivot, const double* fac_row,
                        const double* fac_col, const int* lm2l, const int* rows, const double* rhs,
                        double* sol, int L, int n, int kl, int ku, const int* when, void* stream);

/* Sizes of the argument structs as compiled (binding checks). */
int mh_struct_sizes(int* out, int n);

#ifdef __cplusplus
}
#endif

#endif
