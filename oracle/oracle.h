/*
 * oracle.h -- FP64 CPU restatement of cpsFDTD's 3D time-stepping loop.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the correctness oracle for the
 * HIP engine in cpsfdtd_amd/.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load it, and only as the checker.  The
 * product never links or calls anything here.
 *
 * Layout follows the reference (Fields.cpp:99-131): every 3-D array is
 * (Nx+1)(Ny+1)(Nz+1) doubles, [i][j][k] with k fastest, flattened.
 *
 * Each function in oracle.c cites the reference file:line it restates.
 */
#ifndef CPSFDTD_ORACLE_H
#define CPSFDTD_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

/* Behaviour switches for the reference defects that change results.
 * 0 = behave as the reference source does (the parity target). */
#define ORC_FIX_CPML_HSIGN   1u  /* CPML.cpp:643-682 H-side y/z psi sign (Roden-Gedney when set) */
#define ORC_FIX_TFSF_FARFACE 2u  /* TF_SF.cpp:1222,1255,1288 far-face H written at +1 (index +0 when set) */
#define ORC_SRC_REF_PULSE    4u  /* point sources use Sources.cpp:72-76 Pulse() instead of FWHM/delay */
/* Reproduce the reference as it runs on zero-filled heap memory: the unset
 * CPML H coefficient (CPML.cpp:210) is 0 and the x-face TF/SF H edge loops
 * (TF_SF.cpp:1244-1267), which read unset point fields, add nothing.  Used only
 * to pin the oracle bit-for-bit against the compiled reference. */
#define ORC_REF_ZEROHEAP     8u
/* TF_SF.cpp:155-887 takes every incident value at the cell corner (i, j, k),
 * shifted half a cell along the face normal only; set: at the Yee position of
 * the field component that supplies it (plus half a cell along the corrected
 * E component's own axis, for the E and the H correction of that pair). */
#define ORC_FIX_TFSF_YEE    16u
/* A planted mistake, for tests that must fail on it: the x-slab CPML H terms
 * (CPML.cpp:618-620) applied with the opposite sign.  No run of the
 * reference or the engine uses it. */
#define ORC_TEST_FLIP_HX  (1u << 30)

#define ORC_PORT_PLANE  0
#define ORC_PORT_SPHERE 1

typedef struct {
    const char *material;   /* material.txt   (Structure.cpp:26-52), required   */
    const char *structure;  /* structure.txt  (Structure.cpp:68-157), required  */
    const char *pml;        /* CFS_PML.txt    (CPML.cpp:29-68), NULL = none     */
    const char *tfsf;       /* TF_SF.txt      (TF_SF.cpp:29-959), NULL = none   */
    const char *dl;         /* DrudeLorentz.txt (3-D layout), NULL = none       */
    const char *gain;       /* four-level gain file, NULL = none                */
    const char *source;     /* source.txt (Sources.cpp:29-57), NULL = none      */
    int n_ports;
    const char *const *port_files;
    const int *port_kinds;  /* ORC_PORT_PLANE / ORC_PORT_SPHERE                 */
    int n_src;              /* point sources, soft, added after TF/SF-E         */
    const int *src_ijkc;    /* 4 ints each: i j k comp(0..2 = Ex,Ey,Ez)         */
    const double *src_amp;
    const double *src_phase;
    int nt;                 /* steps; <= 0 takes Nt from TF_SF.txt              */
    int n_freq;             /* port DFT frequencies (Hz)                        */
    const double *freqs;
    int dft_stride;         /* accumulate every dft_stride-th step (>=1)        */
    int record_ports;       /* keep every step's 6-component port samples       */
    int periodic;           /* bit0: x periodic, bit1: y periodic               */
    unsigned flags;         /* ORC_* switches                                   */
} orc_config;

typedef struct orc_sim orc_sim;

/* Returns NULL on failure; orc_last_error() then describes it. */
orc_sim *orc_create(const orc_config *cfg);
void orc_destroy(orc_sim *s);
const char *orc_last_error(void);

/* Geometry / tables. */
void orc_dims(const orc_sim *s, int *nxyz, double *dxyzt);
int orc_nmedia(const orc_sim *s);
void orc_get_id(const orc_sim *s, short *out);           /* (Nx+1)(Ny+1)(Nz+1) */
/* which: 0..3 aE bE aH bH (NoMedia each) */
void orc_get_coef(const orc_sim *s, int which, double *out);
/* stretch tables after CPML kappa folding: 0..5 cExy cEyx cEzx cExz cEzy cEyz,
 * 6..11 cHxy cHyx cHzx cHxz cHzy cHyz; returns length */
int orc_get_stretch(const orc_sim *s, int which, double *out);
/* CPML b/c: which = axis*4 + side*2 + (0:E,1:H); returns length; b,c may be NULL */
int orc_get_cpml(const orc_sim *s, int which, double *b, double *c);
/* TF/SF face points in construction order (TF_SF.cpp:147-887).
 * Returns count.  ijk: i j k part per point (4 ints); tab: dE pE dH pH of
 * component pair 1, then of pair 2, per point (8 doubles). */
int orc_get_tfsf_points(const orc_sim *s, int *ijk, double *tab);
int orc_get_tfsf_wave(const orc_sim *s, double *einc, double *hinc); /* returns Nt+1 */
void orc_get_tfsf_vec(const orc_sim *s, double *e0h0_ce_ch); /* 12 doubles */
/* Ports. */
int orc_port_npts(const orc_sim *s, int port);
void orc_port_points(const orc_sim *s, int port, int *ijk);
/* step-major samples [nt][npts][6] (Ex Ey Ez Hx Hy Hz) when record_ports */
void orc_port_samples(const orc_sim *s, int port, double *out);
/* [npts][n_freq][6][2] (re, im) */
void orc_port_dft(const orc_sim *s, int port, double *out);
/* per-step sum over points of Ex^2+Ey^2+Ez^2 [nt] */
void orc_port_intensity(const orc_sim *s, int port, double *out);

/* Time stepping: runs steps n0 .. n0+count-1 (step index = TF/SF time index). */
int orc_step(orc_sim *s, int count);
int orc_steps_done(const orc_sim *s);
/* comp 0..5 = Ex Ey Ez Hx Hy Hz */
void orc_get_field(const orc_sim *s, int comp, double *out);
void orc_set_field(orc_sim *s, int comp, const double *in);
/* Gain populations (N0..N3 per gain cell, cell order i,j,k) and DL cell count. */
int orc_gain_ncells(const orc_sim *s);
void orc_gain_pop(const orc_sim *s, double *out);
int orc_dl_ncells(const orc_sim *s);
/* TF/SF face variants: with n > 0 every later step applies the listed faces'
 * TF_SF::{x,y,z}{1,2}UpdateHField, then after the E update their
 * Update*EField forms, in that order, in place of UpdateHField/UpdateEField
 * (face = 2 * axis + high: 0 x1, 1 x2, 2 y1, 3 y2, 4 z1, 5 z2; the z1 forms
 * without the profile factor, TF_SF.cpp:1126-1143, 1375-1392).  n = 0
 * restores the whole update. */
int orc_set_tfsf_faces(orc_sim *s, const int *faces, int n);
/* Drude-Lorentz state [cell][3][nD + 2 nL + 1] (JD, JL, JL_1, E_1 per
 * component) and gain polarisation [cell][6] (P x y z, previous P x y z);
 * both return the cell count (out may be NULL). */
int orc_dl_state(const orc_sim *s, double *out);
void orc_dl_poles(const orc_sim *s, int *nD_nL);
int orc_gain_state(const orc_sim *s, double *out);
/* i j k of each Drude-Lorentz (kind 0) or gain (kind 1) cell */
int orc_special_cells(const orc_sim *s, int kind, int *ijk);
/* Point source q's waveform, one value per time index 0..nt (added after the
 * E update of step n); returns nt + 1, or 0 for no such source. */
int orc_source_wave(const orc_sim *s, int q, double *out);
/* DL coefficients: C1 C2 C3 then C4[nD] C5[nL] C6[nL] kappa[nD] beta[nD] alpha[nL] xi[nL] gammaL[nL] */
int orc_dl_coefs(const orc_sim *s, double *out);

/* Zero-dimensional single-cell Drude-Lorentz recurrence driven by a prescribed
 * curl sequence (used by the closed-form susceptibility test). */
int orc_dl_0d(const orc_sim *s, int nsteps, const double *curl, double *e_out);
/* Zero-dimensional four-level cell driven by a prescribed field. */
int orc_gain_0d(const orc_sim *s, int nsteps, const double *efield, double *pop_out, double *p_out);

#ifdef __cplusplus
}
#endif
#endif
