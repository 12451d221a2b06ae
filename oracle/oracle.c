/*
 * oracle.c -- FP64 CPU restatement of cpsFDTD's 3-D time-stepping loop.
 *
 * TEST INFRASTRUCTURE ONLY (see oracle.h).  Written from the reference's
 * equations; every block cites the reference file:line it restates.
 * Arithmetic expressions keep the reference's association order so that the
 * integer ID map and every FP64 table are bit-identical to what the
 * reference computes, and so the HIP engine's FP64 form can be compared at
 * the 1e-12 level (in practice bit-for-bit).
 *
 * Decisions on reference defects (identifiers as in the defect register):
 *   F1  H_pml_Coef leaves b,c[npml-1] unset        -> set (loop i<npml)
 *   F2  second-end kappa folding bound Npml?1      -> Npml?2
 *   F4  TF_SF k0 unset, z0 overwritten             -> k0=zs/dz, z0 = file line 3
 *   F6  CPML H-side y/z psi sign                   -> reference sign unless ORC_FIX_CPML_HSIGN
 *   F9  Drude-Lorentz C1 sign                      -> derived (minus) form
 *   F10 Drude-Lorentz C5 missing dt                -> (1+alpha) dt / den
 *   F11 port point count +1 (unset point)          -> extra point dropped
 *   F13 N=(int)(L/d) truncation                    -> kept
 *   F15 TF/SF far-face H written at index+1        -> reference index unless ORC_FIX_TFSF_FARFACE
 *   F16 TF/SF x-face H edge loops read the unset   -> corrected (Hz with the Ey-edge table at
 *       table half and update the wrong component      k=k2, Hy with the Ez-edge table at j=j2)
 *   F18 TF/SF incident values at the cell corner,  -> reference positions unless ORC_FIX_TFSF_YEE
 *       not at each component's Yee position
 * Shapes that reach outside the grid are clipped (the reference writes out of
 * bounds, Structure.cpp:166,231).
 */
#include "oracle.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <ctype.h>

/* Constants.h:19-26 */
static const double pi = 3.1415926535897932384626433832795;
static const double C0 = 2.99792458E8;
static const double mu0 = 1.2566370614E-6;
static const double eps0 = 8.8541878176E-12;
static const double hPlanck = 6.62607004081E-34;

static char g_err[512];
const char *orc_last_error(void) { return g_err; }
#define FAIL(...) do { snprintf(g_err, sizeof g_err, __VA_ARGS__); goto fail; } while (0)

/* ------------------------------------------------------------------------ */
/* line/token reader: istringstream >> semantics over getline()'d lines      */
/* ------------------------------------------------------------------------ */
typedef struct { char *buf; char **line; int n, cur; } lines_t;

static int lines_read(const char *path, lines_t *L)
{
    memset(L, 0, sizeof *L);
    FILE *f = fopen(path, "rb");
    if (!f) return -1;
    fseek(f, 0, SEEK_END);
    long sz = ftell(f);
    fseek(f, 0, SEEK_SET);
    L->buf = (char *)malloc((size_t)sz + 2);
    size_t got = fread(L->buf, 1, (size_t)sz, f);
    fclose(f);
    L->buf[got] = 0;
    int cap = 16;
    L->line = (char **)malloc(sizeof(char *) * cap);
    char *p = L->buf;
    while (*p) {
        if (L->n == cap) { cap *= 2; L->line = (char **)realloc(L->line, sizeof(char *) * cap); }
        L->line[L->n++] = p;
        char *e = strchr(p, '\n');
        if (!e) break;
        *e = 0;
        p = e + 1;
    }
    return 0;
}
static void lines_free(lines_t *L) { free(L->buf); free(L->line); }
/* getline(): past EOF yields an empty line */
static const char *lines_next(lines_t *L) { return L->cur < L->n ? L->line[L->cur++] : ""; }

static int tok(const char **p, char *out, int cap)
{
    const char *s = *p;
    while (*s && isspace((unsigned char)*s)) s++;
    if (!*s) return 0;
    int n = 0;
    while (*s && !isspace((unsigned char)*s)) { if (n < cap - 1) out[n++] = *s; s++; }
    out[n] = 0;
    *p = s;
    return 1;
}
static int tok_d(const char **p, double *v)
{
    char b[128];
    if (!tok(p, b, sizeof b)) return 0;
    char *e;
    *v = strtod(b, &e);
    return e != b;
}
static int tok_i(const char **p, int *v)
{
    char b[128];
    if (!tok(p, b, sizeof b)) return 0;
    char *e;
    long x = strtol(b, &e, 10);
    *v = (int)x;
    return e != b;
}

/* ------------------------------------------------------------------------ */
/* simulation state                                                          */
/* ------------------------------------------------------------------------ */
typedef struct {
    int ok;
    int npml[3][2];
    double m, malp, kapmax, sig_ratio, alpmax;
    double *bE[3][2], *cE[3][2], *bH[3][2], *cH[3][2];
    int lenE[3][2], lenH[3][2];
    /* psi, reference-shaped boxes: E-side (driven by H differences) and H-side */
    double *qE[3][2][2], *qH[3][2][2];
    double cpmlE, cpmlH;
} cpml_t;

/* part: 3 main, 1 comp1, 2 comp2; [c]: incident delay/profile for the E and
 * the H correction of component pair c (TF_SF.cpp's *_E1/_H1, *_E2/_H2) */
typedef struct { int i, j, k, part; double dE[2], pE[2], dH[2], pH[2]; } tfsf_pt;

typedef struct {
    int ok;
    double lambda, dlambda, w0, z0, zR, xs, ys, zs, theta, phi, E0, psiE;
    int i0, j0, k0, n[3][2], i1, i2, j1, j2, k1, k2;
    double E0x, E0y, E0z, H0x, H0y, H0z, kx, ky, kz, kabs, l0;
    double CEx, CEy, CEz, CHx, CHy, CHz;
    int npts, Nt, cwmode, face_start[7];
    tfsf_pt *pt;
    double *Einc, *Hinc;
    int skip_axis[3];
} tfsf_t;

typedef struct {
    int ok, nD, nL, ncell;
    double sigma, epsInf, *wD, *gD, *wL, *dL, *deL;
    double *kap, *beta, *alpha, *xi, *gamL, C1, C2, C3, *C4, *C5, *C6;
    int *cell;        /* flat index */
    double *st;       /* per cell: 3 comps x (nD + 2 nL + 1) */
} dl_t;

typedef struct {
    int ok, ncell;
    double lambda_a, gamma_a, t32, t21, t10, Ntot, Wp;
    double omega_a, kappa_a, CPa1, CPa2, CPa3;
    int *cell;
    double *P, *P1, *Eprev, *N;  /* P,P1,Eprev: 3 per cell; N: 4 per cell (N0..N3) */
} gain_t;

typedef struct {
    int kind, npts;
    int *ijk;
    double *samples;  /* [nt][npts][6] */
    double *dft;      /* [npts][nf][6][2] */
    double *inten;    /* [nt] */
} port_t;

struct orc_sim {
    /* Structure */
    double Lx, Ly, Lz, dx, dy, dz, dt;
    int Nx, Ny, Nz, NoMedia;
    double *epsR, *muR, *sig, *sigM, *aE, *bE, *aH, *bH;
    double *cst[12]; int cstlen[12];
    short *ID;
    double *F[6];
    size_t ncell;
    int periodic, pbx, pby;
    unsigned flags;
    cpml_t pml;
    tfsf_t tf;
    dl_t dl;
    gain_t gn;
    /* sources */
    int nsrc; int *src; double *swave; /* [nsrc][nt+1] */
    /* ports */
    int nports; port_t *port; int nf; double *freqs; int stride, record;
    int nt, ndone;
    double *twE, *twH; /* [nf][nt] cos/sin pairs */
    int nfaces, faces[64]; /* TF/SF face variants in place of the whole update */
};

enum { T_cExy, T_cEyx, T_cEzx, T_cExz, T_cEzy, T_cEyz, T_cHxy, T_cHyx, T_cHzx, T_cHxz, T_cHzy, T_cHyz };

#define IDX(s, i, j, k) ((((size_t)(i)) * (size_t)((s)->Ny + 1) + (size_t)(j)) * (size_t)((s)->Nz + 1) + (size_t)(k))

/* ------------------------------------------------------------------------ */
/* Structure (Structure.cpp)                                                 */
/* ------------------------------------------------------------------------ */
typedef struct {
    int index, material_index;
    char type[32];
    double theta, phi, psi, x0, y0, z0, a, b, c;
    int ix0, jy0, kz0, ia, jb, kc;
} shape_t;

static void setid(orc_sim *s, int i, int j, int k, int m)
{
    if (i < 0 || j < 0 || k < 0 || i > s->Nx || j > s->Ny || k > s->Nz) return; /* clipped */
    s->ID[IDX(s, i, j, k)] = (short)m;
}

/* Structure.cpp:159-203 */
static void paint_cuboid(orc_sim *s, const shape_t *sh)
{
    if (sh->theta == 0 && sh->phi == 0 && sh->psi == 0) {
        for (int i = 0; i < sh->ia; ++i)
            for (int j = 0; j < sh->jb; ++j)
                for (int k = 0; k < sh->kc; ++k)
                    setid(s, i + sh->ix0, j + sh->jy0, k + sh->kz0, sh->material_index);
    } else {
        double x1, y1, z1, x2, y2, z2, x3, y3, z3;
        int i3, j3, k3;
        for (double x = 0; x < sh->a; x = x + (s->dx / 2.0))
            for (double y = 0; y < sh->b; y = y + (s->dy / 2.0))
                for (double z = 0; z < sh->c; z = z + (s->dz / 2.0)) {
                    x1 = x;
                    y1 = cos(sh->theta) * (y) - sin(sh->theta) * (z);
                    z1 = sin(sh->theta) * (y) + cos(sh->theta) * (z);
                    x2 = cos(sh->phi) * (x1) + sin(sh->phi) * (z1);
                    y2 = y1;
                    z2 = -sin(sh->phi) * (x1) + cos(sh->phi) * (z1);
                    x3 = cos(sh->psi) * (x2) - sin(sh->psi) * (y2);
                    y3 = sin(sh->psi) * (x2) + cos(sh->psi) * (y2);
                    z3 = z2;
                    i3 = (int)((x3 + sh->x0) / s->dx);
                    j3 = (int)((y3 + sh->y0) / s->dy);
                    k3 = (int)((z3 + sh->z0) / s->dz);
                    if ((i3 >= 0) && (i3 <= s->Nx) && (j3 >= 0) && (j3 <= s->Ny) && (k3 >= 0) && (k3 <= s->Nz))
                        s->ID[IDX(s, i3, j3, k3)] = (short)sh->material_index;
                }
    }
}

/* Structure.cpp:205-221 */
static void paint_ellipsoid(orc_sim *s, const shape_t *sh)
{
    for (int i = 0; i < s->Nx; i++)
        for (int j = 0; j < s->Ny; j++)
            for (int k = 0; k < s->Nz; k++) {
                double temp = ((double)i * s->dx - sh->x0) * ((double)i * s->dx - sh->x0) / (sh->a * sh->a)
                            + ((double)j * s->dy - sh->y0) * ((double)j * s->dy - sh->y0) / (sh->b * sh->b)
                            + ((double)k * s->dz - sh->z0) * ((double)k * s->dz - sh->z0) / (sh->c * sh->c);
                if (temp <= 1.0) s->ID[IDX(s, i, j, k)] = (short)sh->material_index;
            }
}

/* Structure.cpp:223-236 */
static void paint_cylinder(orc_sim *s, const shape_t *sh)
{
    for (int k = 0; k < sh->kc; k++)
        for (int i = 0; i < s->Nx; i++)
            for (int j = 0; j < s->Ny; j++) {
                double temp = ((double)i * s->dx - sh->x0) * ((double)i * s->dx - sh->x0) / (sh->a * sh->a)
                            + ((double)j * s->dy - sh->y0) * ((double)j * s->dy - sh->y0) / (sh->b * sh->b);
                if (temp <= 1.0) setid(s, i, j, k + sh->kz0, sh->material_index);
            }
}

/* Structure.cpp:238-306 */
static void paint_multi(orc_sim *s, const shape_t *multi, const shape_t *single)
{
    shape_t my = *single;
    my.ia = (int)(my.a / s->dx);
    my.jb = (int)(my.b / s->dy);
    my.kc = (int)(my.c / s->dz);
    for (int i = 0; i < multi->theta; ++i) {
        my.x0 = multi->x0 + i * multi->a;
        my.ix0 = (int)(my.x0 / s->dx);
        for (int j = 0; j < multi->phi; ++j) {
            my.y0 = multi->y0 + j * multi->b;
            my.jy0 = (int)(my.y0 / s->dy);
            for (int k = 0; k < multi->psi; ++k) {
                my.z0 = multi->z0 + k * multi->c;
                my.kz0 = (int)(my.z0 / s->dz);
                if (!strcmp(single->type, "cuboid")) paint_cuboid(s, &my);
                else if (!strcmp(single->type, "ellipsoid")) paint_ellipsoid(s, &my);
                else if (!strcmp(single->type, "cylinder")) paint_cylinder(s, &my);
            }
        }
    }
}

/* Structure.cpp:26-52 (material) and :68-157 (mesh) */
static int load_structure(orc_sim *s, const char *matf, const char *meshf)
{
    lines_t L;
    const char *p;
    shape_t *sh = NULL;
    if (lines_read(matf, &L)) FAIL("cannot read material file %s", matf);
    p = lines_next(&L);
    if (!tok_i(&p, &s->NoMedia) || s->NoMedia <= 0 || s->NoMedia > 255) { lines_free(&L); FAIL("bad NoMedia in %s", matf); }
    s->epsR = calloc(s->NoMedia, sizeof(double)); s->muR = calloc(s->NoMedia, sizeof(double));
    s->sig = calloc(s->NoMedia, sizeof(double)); s->sigM = calloc(s->NoMedia, sizeof(double));
    for (int i = 0; i < s->NoMedia; i++) {
        p = lines_next(&L);
        if (!tok_d(&p, &s->epsR[i]) || !tok_d(&p, &s->muR[i]) || !tok_d(&p, &s->sig[i]) || !tok_d(&p, &s->sigM[i])) {
            lines_free(&L); FAIL("bad material line %d in %s", i + 2, matf);
        }
    }
    lines_free(&L);

    if (lines_read(meshf, &L)) FAIL("cannot read structure file %s", meshf);
    p = lines_next(&L);
    if (!tok_d(&p, &s->Lx) || !tok_d(&p, &s->Ly) || !tok_d(&p, &s->Lz)) { lines_free(&L); FAIL("bad L line"); }
    p = lines_next(&L);
    if (!tok_d(&p, &s->dx) || !tok_d(&p, &s->dy) || !tok_d(&p, &s->dz)) { lines_free(&L); FAIL("bad d line"); }
    int freespace = 0, nsh = 0;
    p = lines_next(&L);
    tok_i(&p, &freespace);
    p = lines_next(&L);
    tok_i(&p, &nsh);
    if (nsh < 0) nsh = 0;
    sh = calloc(nsh > 0 ? nsh : 1, sizeof(shape_t));
    for (int i = 0; i < nsh; i++) {
        p = lines_next(&L);
        tok_i(&p, &sh[i].index);
        tok_i(&p, &sh[i].material_index);
        tok(&p, sh[i].type, sizeof sh[i].type);
        tok_d(&p, &sh[i].theta); tok_d(&p, &sh[i].phi); tok_d(&p, &sh[i].psi);
        tok_d(&p, &sh[i].x0); tok_d(&p, &sh[i].y0); tok_d(&p, &sh[i].z0);
        tok_d(&p, &sh[i].a); tok_d(&p, &sh[i].b); tok_d(&p, &sh[i].c);
        sh[i].ix0 = (int)(sh[i].x0 / s->dx); sh[i].jy0 = (int)(sh[i].y0 / s->dy); sh[i].kz0 = (int)(sh[i].z0 / s->dz);
        sh[i].ia = (int)(sh[i].a / s->dx); sh[i].jb = (int)(sh[i].b / s->dy); sh[i].kc = (int)(sh[i].c / s->dz);
    }
    lines_free(&L);
    s->Nx = (int)(s->Lx / s->dx);
    s->Ny = (int)(s->Ly / s->dy);
    s->Nz = (int)(s->Lz / s->dz);
    if (s->Nx < 2 || s->Ny < 2 || s->Nz < 2) FAIL("grid too small");
    s->dt = 0.99 / (C0 * sqrt(1.0 / (s->dx * s->dx) + 1.0 / (s->dy * s->dy) + 1.0 / (s->dz * s->dz)));
    s->ncell = (size_t)(s->Nx + 1) * (s->Ny + 1) * (s->Nz + 1);
    s->ID = calloc(s->ncell, sizeof(short));
    if (!freespace) {
        for (int i = 0; i < nsh; i++) {
            if (!strcmp(sh[i].type, "cuboid")) paint_cuboid(s, &sh[i]);
            else if (!strcmp(sh[i].type, "ellipsoid")) paint_ellipsoid(s, &sh[i]);
            else if (!strcmp(sh[i].type, "cylinder")) paint_cylinder(s, &sh[i]);
            else if (!strcmp(sh[i].type, "multi")) {
                int t = sh[i].material_index;
                if (t < 0 || t >= nsh) FAIL("multi shape template index %d out of range", t);
                paint_multi(s, &sh[i], &sh[t]);
            }
        }
    }
    for (size_t c = 0; c < s->ncell; c++)
        if (s->ID[c] < 0 || s->ID[c] >= s->NoMedia) FAIL("material index %d not defined in material file", s->ID[c]);
    free(sh);
    return 0;
fail:
    free(sh);
    return -1;
}

/* Structure.cpp:309-359 */
static void init_coefficients(orc_sim *s)
{
    int n = s->NoMedia;
    s->aE = calloc(n, sizeof(double)); s->bE = calloc(n, sizeof(double));
    s->aH = calloc(n, sizeof(double)); s->bH = calloc(n, sizeof(double));
    int len[12] = {s->Nz, s->Nz, s->Ny, s->Ny, s->Nx, s->Nx, s->Nz, s->Nz, s->Ny, s->Ny, s->Nx, s->Nx};
    double inv[12] = {1.0 / s->dz, 1.0 / s->dz, 1.0 / s->dy, 1.0 / s->dy, 1.0 / s->dx, 1.0 / s->dx,
                      1.0 / s->dz, 1.0 / s->dz, 1.0 / s->dy, 1.0 / s->dy, 1.0 / s->dx, 1.0 / s->dx};
    for (int t = 0; t < 12; t++) {
        s->cstlen[t] = len[t];
        s->cst[t] = malloc(sizeof(double) * len[t]);
        for (int i = 0; i < len[t]; i++) s->cst[t][i] = inv[t];
    }
    double dt = s->dt;
    for (int i = 0; i < n; ++i) {
        s->aE[i] = (2.0 * eps0 * s->epsR[i] - s->sig[i] * dt) / (2.0 * eps0 * s->epsR[i] + s->sig[i] * dt);
        s->bE[i] = (2.0 * dt) / (2.0 * eps0 * s->epsR[i] + s->sig[i] * dt);
        s->aH[i] = (2.0 * mu0 * s->muR[i] - s->sigM[i] * dt) / (2.0 * mu0 * s->muR[i] + s->sigM[i] * dt);
        s->bH[i] = (2.0 * dt) / (2.0 * mu0 * s->muR[i] + s->sigM[i] * dt);
    }
}

/* ------------------------------------------------------------------------ */
/* CPML (CPML.cpp)                                                           */
/* ------------------------------------------------------------------------ */
/* CPML.cpp:139-183 */
static void e_pml_coef(double *kappa, double *b, double *c, int firstend, int npml, double dt, double del,
                       double m, double ma, double sigmax, double kapmax, double alpmax)
{
    double *sigma = malloc(sizeof(double) * (npml + 1)), *alpha = malloc(sizeof(double) * (npml + 1));
    for (int i = 1; i <= npml + 1; ++i) {
        double psk, pa;
        if (firstend) {
            psk = pow((((double)npml + 1.0 - i) / ((double)npml)), m);
            pa = pow((((double)i - 1.0) / ((double)npml)), ma);
        } else {
            psk = pow((((double)i - 1.0) / npml), m);
            pa = pow((((double)npml + 1.0 - i) / npml), ma);
        }
        sigma[i - 1] = sigmax * psk;
        kappa[i - 1] = 1.0 + (kapmax - 1.0) * psk;
        alpha[i - 1] = alpmax * pa;
    }
    for (int i = 0; i <= npml; ++i) {
        b[i] = exp(-(sigma[i] / kappa[i] + alpha[i]) * dt / eps0);
        if ((sigma[i] == 0.0) && (alpha[i] == 0.0)) c[i] = 0.0;
        else c[i] = sigma[i] * (b[i] - 1.0) / (sigma[i] + kappa[i] * alpha[i]) / kappa[i] / del;
    }
    free(sigma); free(alpha);
}

/* CPML.cpp:185-229, loop bound corrected to npml (F1); refzero keeps the
 * reference's unset last entry as zero-filled memory would leave it. */
static void h_pml_coef(double *kappa, double *b, double *c, int firstend, int npml, double dt, double del,
                       double m, double ma, double sigmax, double kapmax, double alpmax)
{
    double *sigma = malloc(sizeof(double) * npml), *alpha = malloc(sizeof(double) * npml);
    for (int i = 1; i <= npml; ++i) {
        double psk, pa;
        if (firstend) {
            psk = pow(((npml - i + 0.5) / npml), m);
            pa = pow(((i - 0.5) / npml), ma);
        } else {
            psk = pow(((i - 0.5) / npml), m);
            pa = pow(((npml - i + 0.5) / npml), ma);
        }
        sigma[i - 1] = sigmax * psk;
        kappa[i - 1] = 1.0 + (kapmax - 1.0) * psk;
        alpha[i - 1] = alpmax * pa;
    }
    for (int i = 0; i < npml; ++i) {
        b[i] = exp(-(sigma[i] / kappa[i] + alpha[i]) * dt / eps0);
        if ((sigma[i] == 0.0) && (alpha[i] == 0.0)) c[i] = 0.0;
        else c[i] = sigma[i] * (b[i] - 1.0) / (sigma[i] + kappa[i] * alpha[i]) / kappa[i] / del;
    }
    free(sigma); free(alpha);
}

/* CPML.cpp:29-68 (parse) and :231-527 (init) */
static int load_cpml(orc_sim *s, const char *path)
{
    cpml_t *P = &s->pml;
    lines_t L;
    const char *p;
    if (lines_read(path, &L)) FAIL("cannot read pml file %s", path);
    for (int a = 0; a < 3; a++) {
        p = lines_next(&L);
        if (!tok_i(&p, &P->npml[a][0]) || !tok_i(&p, &P->npml[a][1])) { lines_free(&L); FAIL("bad pml line %d", a + 1); }
    }
    p = lines_next(&L);
    if (!tok_d(&p, &P->m) || !tok_d(&p, &P->malp)) { lines_free(&L); FAIL("bad pml line 4"); }
    p = lines_next(&L);
    if (!tok_d(&p, &P->kapmax) || !tok_d(&p, &P->sig_ratio)) { lines_free(&L); FAIL("bad pml line 5"); }
    p = lines_next(&L);
    if (!tok_d(&p, &P->alpmax)) { lines_free(&L); FAIL("bad pml line 6"); }
    lines_free(&L);
    if (s->pbx) P->npml[0][0] = P->npml[0][1] = 0;
    if (s->pby) P->npml[1][0] = P->npml[1][1] = 0;
    int N[3] = {s->Nx, s->Ny, s->Nz};
    double d[3] = {s->dx, s->dy, s->dz};
    for (int a = 0; a < 3; a++)
        for (int e = 0; e < 2; e++)
            if (P->npml[a][e] < 0 || P->npml[a][0] + P->npml[a][1] >= N[a]) FAIL("pml thickness out of range");
    double eta0 = sqrt(mu0 / eps0);
    /* kappa folding targets: E side (axis) -> two E tables; H side -> two H tables */
    const int foldE[3][2] = {{T_cEzy, T_cEyz}, {T_cEzx, T_cExz}, {T_cExy, T_cEyx}};
    const int foldH[3][2] = {{T_cHzy, T_cHyz}, {T_cHzx, T_cHxz}, {T_cHxy, T_cHyx}};
    for (int a = 0; a < 3; a++) {
        double sigmax = P->sig_ratio * (0.8 * (P->m + 1) / (d[a] * eta0));
        for (int e = 0; e < 2; e++) {
            int n = P->npml[a][e];
            if (n == 0) continue;  /* no slab on this face */
            double *kap = malloc(sizeof(double) * (n + 1));
            P->lenE[a][e] = n + 1;
            P->bE[a][e] = malloc(sizeof(double) * (n + 1));
            P->cE[a][e] = malloc(sizeof(double) * (n + 1));
            e_pml_coef(kap, P->bE[a][e], P->cE[a][e], e == 0, n, s->dt, d[a], P->m, P->malp, sigmax, P->kapmax, P->alpmax);
            if (e == 0) {
                for (int i = 0; i <= n; ++i) { s->cst[foldE[a][0]][i] /= kap[i]; s->cst[foldE[a][1]][i] /= kap[i]; }
            } else {
                for (int i = 0; i < n; ++i) {
                    s->cst[foldE[a][0]][N[a] - n + i] /= kap[i];
                    s->cst[foldE[a][1]][N[a] - n + i] /= kap[i];
                }
            }
            free(kap);
        }
    }
    for (int a = 0; a < 3; a++) {
        double sigmax = P->sig_ratio * (0.8 * (P->m + 1) / (d[a] * eta0));
        for (int e = 0; e < 2; e++) {
            int n = P->npml[a][e];
            if (n == 0) continue;
            double *kap = malloc(sizeof(double) * n);
            P->lenH[a][e] = n;
            P->bH[a][e] = calloc(n, sizeof(double));
            P->cH[a][e] = calloc(n, sizeof(double));
            h_pml_coef(kap, P->bH[a][e], P->cH[a][e], e == 0, n, s->dt, d[a], P->m, P->malp, sigmax, P->kapmax, P->alpmax);
            if (s->flags & ORC_REF_ZEROHEAP) { P->bH[a][e][n - 1] = 0.0; P->cH[a][e][n - 1] = 0.0; }
            for (int i = 0; i < n; ++i) {
                int t = (e == 0) ? i : N[a] - n + i;
                s->cst[foldH[a][0]][t] /= kap[i];
                s->cst[foldH[a][1]][t] /= kap[i];
            }
            free(kap);
        }
    }
    P->cpmlE = s->dt / eps0;
    P->cpmlH = s->bH[0];
    /* psi boxes, shaped as CPML.cpp:235-350 */
    for (int a = 0; a < 3; a++)
        for (int e = 0; e < 2; e++) {
            int n = P->npml[a][e];
            if (!n) continue;
            size_t sz = (size_t)(n + 1);
            for (int b = 0; b < 3; b++) if (b != a) sz *= (size_t)(N[b] + 1);
            for (int q = 0; q < 2; q++) { P->qE[a][e][q] = calloc(sz, sizeof(double)); P->qH[a][e][q] = calloc(sz, sizeof(double)); }
        }
    P->ok = 1;
    return 0;
fail:
    return -1;
}

/* psi box index: slab-local coordinate l along axis a (box thickness n+1),
 * full range on the other two axes */
static inline size_t qidx(const orc_sim *s, int a, int n, int l, int i, int j, int k)
{
    size_t n1 = (size_t)s->Ny + 1, n2 = (size_t)s->Nz + 1;
    if (a == 0) return ((size_t)l * n1 + j) * n2 + k;
    if (a == 1) return ((size_t)i * (n + 1) + l) * n2 + k;
    return ((size_t)i * n1 + j) * (size_t)(n + 1) + l;
}

/* CPML.cpp:529-607 */
static void cpml_update_E(orc_sim *s)
{
    cpml_t *P = &s->pml;
    double *Ex = s->F[0], *Ey = s->F[1], *Ez = s->F[2], *Hx = s->F[3], *Hy = s->F[4], *Hz = s->F[5];
    int Nx = s->Nx, Ny = s->Ny, Nz = s->Nz;
    int ie0 = s->pbx ? 0 : 1, je0 = s->pby ? 0 : 1;
    double cE = P->cpmlE;
    /* x slabs: Ey -= psi(dHz/dx), Ez += psi(dHy/dx) */
    for (int e = 0; e < 2; e++) {
        int n = P->npml[0][e];
        if (!n) continue;
        for (int l = 0; l < n; l++) {
            int i = e == 0 ? l + 1 : Nx - n + l;
            int ci = e == 0 ? l + 1 : l;
            double b = P->bE[0][e][ci], c = P->cE[0][e][ci];
            double *q1 = P->qE[0][e][0], *q2 = P->qE[0][e][1];
#pragma omp parallel for
            for (int j = je0; j < Ny; ++j)
                for (int k = 1; k < Nz; ++k) {
                    size_t p = IDX(s, i, j, k), qi = qidx(s, 0, n, ci, i, j, k);
                    q1[qi] = b * q1[qi] + c * (Hz[p] - Hz[IDX(s, i - 1, j, k)]);
                    Ey[p] = Ey[p] - q1[qi] * cE;
                    q2[qi] = b * q2[qi] + c * (Hy[p] - Hy[IDX(s, i - 1, j, k)]);
                    Ez[p] = Ez[p] + q2[qi] * cE;
                }
        }
    }
    /* y slabs: Ez -= psi(dHx/dy), Ex += psi(dHz/dy) */
    for (int e = 0; e < 2; e++) {
        int n = P->npml[1][e];
        if (!n) continue;
        double *q1 = P->qE[1][e][0], *q2 = P->qE[1][e][1];
#pragma omp parallel for
        for (int i = ie0; i < Nx; ++i)
            for (int l = 0; l < n; l++) {
                int j = e == 0 ? l + 1 : Ny - n + l;
                int cj = e == 0 ? l + 1 : l;
                double b = P->bE[1][e][cj], c = P->cE[1][e][cj];
                int jm = j - 1;
                for (int k = 1; k < Nz; ++k) {
                    size_t p = IDX(s, i, j, k), qi = qidx(s, 1, n, cj, i, j, k);
                    q1[qi] = b * q1[qi] + c * (Hx[p] - Hx[IDX(s, i, jm, k)]);
                    Ez[p] = Ez[p] - q1[qi] * cE;
                    q2[qi] = b * q2[qi] + c * (Hz[p] - Hz[IDX(s, i, jm, k)]);
                    Ex[p] = Ex[p] + q2[qi] * cE;
                }
            }
    }
    /* z slabs: Ex -= psi(dHy/dz), Ey += psi(dHx/dz) */
    for (int e = 0; e < 2; e++) {
        int n = P->npml[2][e];
        if (!n) continue;
        double *q1 = P->qE[2][e][0], *q2 = P->qE[2][e][1];
#pragma omp parallel for
        for (int i = ie0; i < Nx; ++i)
            for (int j = je0; j < Ny; ++j)
                for (int l = 0; l < n; l++) {
                    int k = e == 0 ? l + 1 : Nz - n + l;
                    int ck = e == 0 ? l + 1 : l;
                    double b = P->bE[2][e][ck], c = P->cE[2][e][ck];
                    size_t p = IDX(s, i, j, k), qi = qidx(s, 2, n, ck, i, j, k);
                    q1[qi] = b * q1[qi] + c * (Hy[p] - Hy[p - 1]);
                    Ex[p] = Ex[p] - q1[qi] * cE;
                    q2[qi] = b * q2[qi] + c * (Hx[p] - Hx[p - 1]);
                    Ey[p] = Ey[p] + q2[qi] * cE;
                }
    }
}

/* CPML.cpp:609-689; ORC_FIX_CPML_HSIGN flips the y/z-slab signs (F6) */
static void cpml_update_H(orc_sim *s)
{
    cpml_t *P = &s->pml;
    double *Ex = s->F[0], *Ey = s->F[1], *Ez = s->F[2], *Hx = s->F[3], *Hy = s->F[4], *Hz = s->F[5];
    int Nx = s->Nx, Ny = s->Ny, Nz = s->Nz;
    double cH = P->cpmlH;
    double sg = (s->flags & ORC_FIX_CPML_HSIGN) ? -1.0 : 1.0;
    const int flipx = (s->flags & ORC_TEST_FLIP_HX) != 0;
    /* x slabs: Hy += psi(dEz/dx), Hz -= psi(dEy/dx) */
    for (int e = 0; e < 2; e++) {
        int n = P->npml[0][e];
        if (!n) continue;
        for (int l = 0; l < n; l++) {
            int i = e == 0 ? l : Nx - n + l;
            double b = P->bH[0][e][l], c = P->cH[0][e][l];
            double *q1 = P->qH[0][e][0], *q2 = P->qH[0][e][1];
#pragma omp parallel for
            for (int j = 0; j < Ny; ++j)
                for (int k = 0; k < Nz; ++k) {
                    size_t p = IDX(s, i, j, k), qi = qidx(s, 0, n, l, i, j, k);
                    q1[qi] = b * q1[qi] + c * (Ez[IDX(s, i + 1, j, k)] - Ez[p]);
                    q2[qi] = b * q2[qi] + c * (Ey[IDX(s, i + 1, j, k)] - Ey[p]);
                    if (flipx) {  /* planted mistake (ORC_TEST_FLIP_HX) */
                        Hy[p] = Hy[p] - q1[qi] * cH;
                        Hz[p] = Hz[p] + q2[qi] * cH;
                    } else {
                        Hy[p] = Hy[p] + q1[qi] * cH;
                        Hz[p] = Hz[p] - q2[qi] * cH;
                    }
                }
        }
    }
    /* y slabs (reference): Hz -= psi(dEx/dy), Hx += psi(dEz/dy) */
    for (int e = 0; e < 2; e++) {
        int n = P->npml[1][e];
        if (!n) continue;
        double *q1 = P->qH[1][e][0], *q2 = P->qH[1][e][1];
#pragma omp parallel for
        for (int i = 0; i < Nx; ++i)
            for (int l = 0; l < n; l++) {
                int j = e == 0 ? l : Ny - n + l;
                double b = P->bH[1][e][l], c = P->cH[1][e][l];
                for (int k = 0; k < Nz; ++k) {
                    size_t p = IDX(s, i, j, k), qi = qidx(s, 1, n, l, i, j, k);
                    q1[qi] = b * q1[qi] + c * (Ex[IDX(s, i, j + 1, k)] - Ex[p]);
                    if (sg > 0) Hz[p] = Hz[p] - q1[qi] * cH; else Hz[p] = Hz[p] + q1[qi] * cH;
                    q2[qi] = b * q2[qi] + c * (Ez[IDX(s, i, j + 1, k)] - Ez[p]);
                    if (sg > 0) Hx[p] = Hx[p] + q2[qi] * cH; else Hx[p] = Hx[p] - q2[qi] * cH;
                }
            }
    }
    /* z slabs (reference): Hx -= psi(dEy/dz), Hy += psi(dEx/dz) */
    for (int e = 0; e < 2; e++) {
        int n = P->npml[2][e];
        if (!n) continue;
        double *q1 = P->qH[2][e][0], *q2 = P->qH[2][e][1];
#pragma omp parallel for
        for (int i = 0; i < Nx; ++i)
            for (int j = 0; j < Ny; ++j)
                for (int l = 0; l < n; l++) {
                    int k = e == 0 ? l : Nz - n + l;
                    double b = P->bH[2][e][l], c = P->cH[2][e][l];
                    size_t p = IDX(s, i, j, k), qi = qidx(s, 2, n, l, i, j, k);
                    q1[qi] = b * q1[qi] + c * (Ey[p + 1] - Ey[p]);
                    if (sg > 0) Hx[p] = Hx[p] - q1[qi] * cH; else Hx[p] = Hx[p] + q1[qi] * cH;
                    q2[qi] = b * q2[qi] + c * (Ex[p + 1] - Ex[p]);
                    if (sg > 0) Hy[p] = Hy[p] + q2[qi] * cH; else Hy[p] = Hy[p] - q2[qi] * cH;
                }
    }
}

/* ------------------------------------------------------------------------ */
/* BasicFDTD.cpp                                                             */
/* ------------------------------------------------------------------------ */
/* BasicFDTD.cpp:67-81, with periodic wrap of the +1 neighbour when enabled */
static void basic_H(orc_sim *s)
{
    double *Ex = s->F[0], *Ey = s->F[1], *Ez = s->F[2], *Hx = s->F[3], *Hy = s->F[4], *Hz = s->F[5];
    int Nx = s->Nx, Ny = s->Ny, Nz = s->Nz;
    const double *cHxy = s->cst[T_cHxy], *cHyx = s->cst[T_cHyx], *cHzx = s->cst[T_cHzx];
    const double *cHxz = s->cst[T_cHxz], *cHzy = s->cst[T_cHzy], *cHyz = s->cst[T_cHyz];
#pragma omp parallel for
    for (int i = 0; i < Nx; ++i) {
        int ip = (s->pbx && i + 1 == Nx) ? 0 : i + 1;
        for (int j = 0; j < Ny; ++j) {
            int jp = (s->pby && j + 1 == Ny) ? 0 : j + 1;
            for (int k = 0; k < Nz; ++k) {
                size_t p = IDX(s, i, j, k);
                int id = s->ID[p];
                double aH = s->aH[id], bH = s->bH[id];
                size_t pj = IDX(s, i, jp, k), pi_ = IDX(s, ip, j, k);
                Hx[p] = aH * Hx[p] - bH * (Ez[pj] - Ez[p]) * cHxz[j] + bH * (Ey[p + 1] - Ey[p]) * cHxy[k];
                Hy[p] = aH * Hy[p] - bH * (Ex[p + 1] - Ex[p]) * cHyx[k] + bH * (Ez[pi_] - Ez[p]) * cHyz[i];
                Hz[p] = aH * Hz[p] + bH * (Ex[pj] - Ex[p]) * cHzx[j] - bH * (Ey[pi_] - Ey[p]) * cHzy[i];
            }
        }
    }
}

/* Drude-Lorentz ADE for one cell (3-D form of Drude_Lorentz.cpp:153-227) */
static void dl_cell(orc_sim *s, int c, size_t p, size_t pim, size_t pjm)
{
    dl_t *D = &s->dl;
    double *Ex = s->F[0], *Ey = s->F[1], *Ez = s->F[2], *Hx = s->F[3], *Hy = s->F[4], *Hz = s->F[5];
    int nD = D->nD, nL = D->nL, per = nD + 2 * nL + 1;
    int i = (int)(p / ((size_t)(s->Ny + 1) * (s->Nz + 1)));
    int j = (int)((p / (size_t)(s->Nz + 1)) % (size_t)(s->Ny + 1));
    int k = (int)(p % (size_t)(s->Nz + 1));
    double dA[3], dB[3];
    dA[0] = (Hz[p] - Hz[pjm]); dB[0] = (Hy[p] - Hy[p - 1]);
    dA[1] = (Hx[p] - Hx[p - 1]);      dB[1] = (Hz[p] - Hz[pim]);
    dA[2] = (Hy[p] - Hy[pim]);        dB[2] = (Hx[p] - Hx[pjm]);
    double cA[3] = {s->cst[T_cExz][j], s->cst[T_cEyx][k], s->cst[T_cEzy][i]};
    double cB[3] = {s->cst[T_cExy][k], s->cst[T_cEyz][i], s->cst[T_cEzx][j]};
    double *Fc[3] = {Ex, Ey, Ez};
    for (int q = 0; q < 3; q++) {
        double *st = D->st + ((size_t)c * 3 + q) * per;
        double *JD = st, *JL = st + nD, *JL1 = st + nD + nL, *E1 = st + nD + 2 * nL;
        double temp = Fc[q][p];
        double t1 = 0.0, t2 = 0.0;
        for (int pp = 0; pp < nD; pp++) t1 = t1 + D->C4[pp] * JD[pp];
        for (int pp = 0; pp < nL; pp++) t2 = t2 + D->C5[pp] * JL[pp] + D->C6[pp] * JL1[pp];
        double En = D->C1 * Fc[q][p] + D->C2 * (*E1) + D->C3 * dA[q] * cA[q] - D->C3 * dB[q] * cB[q] - t1 - t2;
        Fc[q][p] = En;
        for (int pp = 0; pp < nD; pp++) JD[pp] = D->kap[pp] * JD[pp] + D->beta[pp] * (En + temp);
        for (int pp = 0; pp < nL; pp++) {
            double t0 = JL[pp];
            JL[pp] = D->alpha[pp] * JL[pp] + D->xi[pp] * JL1[pp] + D->gamL[pp] * (En - *E1) / 2.0 / s->dt;
            JL1[pp] = t0;
        }
        *E1 = temp;
    }
}

/* Four-level gain cell, polarisation part (3-D model, see load_gain) */
static void gain_pol(orc_sim *s, int c, size_t p)
{
    gain_t *G = &s->gn;
    double *N = G->N + (size_t)c * 4;
    double dN = N[1] - N[2];
    for (int q = 0; q < 3; q++) {
        double En = s->F[q][p];
        double *P = G->P + (size_t)c * 3 + q, *P1 = G->P1 + (size_t)c * 3 + q;
        double Pn = *P;
        *P = G->CPa1 * Pn + G->CPa2 * (*P1) + G->CPa3 * dN * En;
        *P1 = Pn;
        G->Eprev[(size_t)c * 3 + q] = En;
    }
}
/* E -= (bE/dt)(P^{n+1}-P^n) and the population update */
static void gain_post(orc_sim *s, int c, size_t p)
{
    gain_t *G = &s->gn;
    int id = s->ID[p];
    double coef = s->bE[id] / s->dt;
    double dot = 0.0;
    for (int q = 0; q < 3; q++) {
        double dP = G->P[(size_t)c * 3 + q] - G->P1[(size_t)c * 3 + q];
        double En = s->F[q][p] - coef * dP;
        s->F[q][p] = En;
        dot = dot + (En + G->Eprev[(size_t)c * 3 + q]) * dP;
    }
    double R = dot / (2.0 * s->dt) / (hPlanck / (2.0 * pi) * G->omega_a);
    double *N = G->N + (size_t)c * 4;
    double dt = s->dt;
    double h32 = dt / (2.0 * G->t32), h21 = dt / (2.0 * G->t21), h10 = dt / (2.0 * G->t10);
    double N3 = (N[3] * (1.0 - h32) + dt * G->Wp * N[0]) / (1.0 + h32);
    double N2 = (N[2] * (1.0 - h21) + h32 * (N[3] + N3) + dt * R) / (1.0 + h21);
    double N1 = (N[1] * (1.0 - h10) + h21 * (N[2] + N2) - dt * R) / (1.0 + h10);
    N[3] = N3; N[2] = N2; N[1] = N1;
    N[0] = G->Ntot - N1 - N2 - N3;
}

/* BasicFDTD.cpp:25-45 plus the dispersive (id 2) and gain (id 5) cells */
static void basic_E(orc_sim *s)
{
    double *Ex = s->F[0], *Ey = s->F[1], *Ez = s->F[2], *Hx = s->F[3], *Hy = s->F[4], *Hz = s->F[5];
    int Nx = s->Nx, Ny = s->Ny, Nz = s->Nz;
    int ie0 = s->pbx ? 0 : 1, je0 = s->pby ? 0 : 1;
    const double *cExy = s->cst[T_cExy], *cEyx = s->cst[T_cEyx], *cEzx = s->cst[T_cEzx];
    const double *cExz = s->cst[T_cExz], *cEzy = s->cst[T_cEzy], *cEyz = s->cst[T_cEyz];
    if (s->gn.ok) {
#pragma omp parallel for
        for (int c = 0; c < s->gn.ncell; c++) gain_pol(s, c, (size_t)s->gn.cell[c]);
    }
#pragma omp parallel for
    for (int i = ie0; i < Nx; ++i) {
        int im = (i == 0) ? Nx - 1 : i - 1;
        for (int j = je0; j < Ny; ++j) {
            int jm = (j == 0) ? Ny - 1 : j - 1;
            for (int k = 1; k < Nz; ++k) {
                size_t p = IDX(s, i, j, k);
                int id = s->ID[p];
                size_t pim = IDX(s, im, j, k), pjm = IDX(s, i, jm, k);
                if (id == 4) {
                    Ex[p] = 0.0; Ey[p] = 0.0; Ez[p] = 0.0;
                } else if (id != 2) {
                    double aE = s->aE[id], bE = s->bE[id];
                    Ex[p] = aE * Ex[p] + bE * (Hz[p] - Hz[pjm]) * cExz[j] - bE * (Hy[p] - Hy[p - 1]) * cExy[k];
                    Ey[p] = aE * Ey[p] + bE * (Hx[p] - Hx[p - 1]) * cEyx[k] - bE * (Hz[p] - Hz[pim]) * cEyz[i];
                    Ez[p] = aE * Ez[p] + bE * (Hy[p] - Hy[pim]) * cEzy[i] - bE * (Hx[p] - Hx[pjm]) * cEzx[j];
                }
            }
        }
    }
    if (s->dl.ok) {
#pragma omp parallel for
        for (int c = 0; c < s->dl.ncell; c++) {
            size_t p = (size_t)s->dl.cell[c];
            int i = (int)(p / ((size_t)(Ny + 1) * (Nz + 1)));
            int j = (int)((p / (size_t)(Nz + 1)) % (size_t)(Ny + 1));
            int k = (int)(p % (size_t)(Nz + 1));
            int im = (i == 0) ? Nx - 1 : i - 1, jm = (j == 0) ? Ny - 1 : j - 1;
            dl_cell(s, c, p, IDX(s, im, j, k), IDX(s, i, jm, k));
        }
    }
    if (s->gn.ok) {
#pragma omp parallel for
        for (int c = 0; c < s->gn.ncell; c++) gain_post(s, c, (size_t)s->gn.cell[c]);
    }
}

/* ------------------------------------------------------------------------ */
/* TF/SF (TF_SF.cpp)                                                         */
/* ------------------------------------------------------------------------ */
/* TF_SF.cpp:158-166 (profile) and :166 (delay) at one position */
static void tfsf_at(const tfsf_t *T, double dt, double x, double y, double z, double *delay, double *prof)
{
    double r2 = x * x + y * y + z * z;
    double rdk = (x * T->kx + y * T->ky + z * T->kz) / T->kabs;
    double rho2 = r2 - rdk * rdk;
    rdk = rdk - T->z0;
    double wz = T->w0 * sqrt(1 + rdk * rdk / (T->zR * T->zR));
    *prof = exp(-rho2 / wz / wz);
    *delay = ((x * T->kx + y * T->ky + z * T->kz) / C0 - T->l0) / dt;
}

static void tfsf_point(orc_sim *s, tfsf_t *T, int i, int j, int k, int part, int axis, double shiftE)
{
    /* the E components a face of this axis corrects (record parts 1, 2) */
    static const int comp[3][2] = {{1, 2}, {0, 2}, {0, 1}};
    tfsf_pt *q = &T->pt[T->npts++];
    q->i = i; q->j = j; q->k = k; q->part = part;
    for (int c = 0; c < 2; c++) {
        /* ORC_FIX_TFSF_YEE: the pair's Yee offset along its E component's axis */
        double o[3] = {0, 0, 0};
        if (s->flags & ORC_FIX_TFSF_YEE) o[comp[axis][c]] = 0.5;
        double x = s->dx * (i - T->i0 + o[0]), y = s->dy * (j - T->j0 + o[1]), z = s->dz * (k - T->k0 + o[2]);
        double xe = x, ye = y, ze = z;
        if (axis == 0) xe = s->dx * (i - T->i0 + shiftE);
        if (axis == 1) ye = s->dy * (j - T->j0 + shiftE);
        if (axis == 2) ze = s->dz * (k - T->k0 + shiftE);
        tfsf_at(T, s->dt, xe, ye, ze, &q->dE[c], &q->pE[c]);
        tfsf_at(T, s->dt, x, y, z, &q->dH[c], &q->pH[c]);
    }
}

/* TF_SF.cpp:29-959 */
static int load_tfsf(orc_sim *s, const char *path, int nt_override)
{
    tfsf_t *T = &s->tf;
    lines_t L;
    const char *p;
    if (lines_read(path, &L)) FAIL("cannot read tfsf file %s", path);
    p = lines_next(&L); if (!tok_d(&p, &T->lambda)) { lines_free(&L); FAIL("tfsf: lambda"); }
    p = lines_next(&L); if (!tok_d(&p, &T->dlambda)) { lines_free(&L); FAIL("tfsf: dlambda"); }
    p = lines_next(&L); if (!tok_d(&p, &T->w0) || !tok_d(&p, &T->z0)) { lines_free(&L); FAIL("tfsf: w0 z0"); }
    T->zR = pi * T->w0 * T->w0 / T->lambda;
    p = lines_next(&L); if (!tok_d(&p, &T->xs) || !tok_d(&p, &T->ys) || !tok_d(&p, &T->zs)) { lines_free(&L); FAIL("tfsf: xs ys zs"); }
    T->i0 = (int)(T->xs / s->dx); T->j0 = (int)(T->ys / s->dy); T->k0 = (int)(T->zs / s->dz);
    p = lines_next(&L); if (!tok_d(&p, &T->theta) || !tok_d(&p, &T->phi)) { lines_free(&L); FAIL("tfsf: angles"); }
    T->theta = T->theta * pi / 180.0; T->phi = T->phi * pi / 180.0;
    p = lines_next(&L); if (!tok_d(&p, &T->E0) || !tok_d(&p, &T->psiE)) { lines_free(&L); FAIL("tfsf: E0 psi"); }
    T->psiE = T->psiE * pi / 180.0;
    for (int a = 0; a < 3; a++) {
        p = lines_next(&L);
        if (!tok_i(&p, &T->n[a][0]) || !tok_i(&p, &T->n[a][1])) { lines_free(&L); FAIL("tfsf: box line %d", 7 + a); }
    }
    p = lines_next(&L); if (!tok_i(&p, &T->Nt)) { lines_free(&L); FAIL("tfsf: Nt"); }
    p = lines_next(&L); tok_i(&p, &T->cwmode);
    lines_free(&L);
    T->skip_axis[0] = s->pbx; T->skip_axis[1] = s->pby; T->skip_axis[2] = 0;
    if (s->pbx) T->n[0][0] = T->n[0][1] = 0;
    if (s->pby) T->n[1][0] = T->n[1][1] = 0;
    T->i1 = T->n[0][0]; T->i2 = s->Nx - T->n[0][1];
    T->j1 = T->n[1][0]; T->j2 = s->Ny - T->n[1][1];
    T->k1 = T->n[2][0]; T->k2 = s->Nz - T->n[2][1];
    if (T->i1 < 1 && !s->pbx) FAIL("tfsf box touches the x boundary");
    if (T->j1 < 1 && !s->pby) FAIL("tfsf box touches the y boundary");
    if (T->k1 < 1 || T->i2 <= T->i1 || T->j2 <= T->j1 || T->k2 <= T->k1 || T->k2 >= s->Nz) FAIL("tfsf box out of range");
    if ((!s->pbx && T->i2 >= s->Nx) || (!s->pby && T->j2 >= s->Ny)) FAIL("tfsf box out of range");
    double eta0 = sqrt(mu0 / eps0);
    double E0 = T->E0, ps = T->psiE, th = T->theta, ph = T->phi;
    /* TF_SF.cpp:112-117 (Taflove) */
    T->E0x = E0 * (cos(ps) * sin(ph) - sin(ps) * cos(th) * cos(ph));
    T->E0y = E0 * (-cos(ps) * cos(ph) - sin(ps) * cos(th) * sin(ph));
    T->E0z = E0 * sin(ps) * sin(th);
    T->H0x = E0 / eta0 * (sin(ps) * sin(ph) + cos(ps) * cos(th) * cos(ph));
    T->H0y = E0 / eta0 * (-sin(ps) * cos(ph) + cos(ps) * cos(th) * sin(ph));
    T->H0z = -E0 / eta0 * cos(ps) * sin(th);
    T->kx = sin(th) * cos(ph); T->ky = sin(th) * sin(ph); T->kz = cos(th);
    T->kabs = sqrt(T->kx * T->kx + T->ky * T->ky + T->kz * T->kz);
    /* TF_SF.cpp:124-138 */
    int ci[2] = {T->i1, T->i2}, cj[2] = {T->j1, T->j2}, ck[2] = {T->k1, T->k2};
    T->l0 = 0;
    for (int q = 0; q < 8; q++) {
        int a = (q >> 2) & 1, b = (q >> 1) & 1, c = q & 1;
        double v = (T->kx * (ci[a] - T->i0) * s->dx + T->ky * (cj[b] - T->j0) * s->dy + T->kz * (ck[c] - T->k0) * s->dz) / C0;
        if (q == 0 || T->l0 > v) T->l0 = v;
    }
    /* TF_SF.cpp:142-887: six faces, each a main loop then its two edge loops */
    int cap = 2 * ((T->i2 - T->i1 + 1) * (T->j2 - T->j1 + 1) + (T->i2 - T->i1 + 1) * (T->k2 - T->k1 + 1) +
                   (T->j2 - T->j1 + 1) * (T->k2 - T->k1 + 1));
    T->pt = calloc(cap, sizeof(tfsf_pt));
    int fs = 0;
    for (int e = 0; e < 2; e++) { /* y-directed, TF_SF.cpp:149-393 */
        T->face_start[fs++] = T->npts;
        if (s->pby) continue;
        int j = e ? T->j2 : T->j1;
        double sh = e ? 0.5 : -0.5;
        for (int i = T->i1; i < T->i2; ++i) for (int k = T->k1; k < T->k2; ++k) tfsf_point(s, T, i, j, k, 3, 1, sh);
        for (int i = T->i1; i < T->i2; ++i) tfsf_point(s, T, i, j, T->k2, 1, 1, sh);
        if (!s->pbx) for (int k = T->k1; k < T->k2; ++k) tfsf_point(s, T, T->i2, j, k, 2, 1, sh);
    }
    for (int e = 0; e < 2; e++) { /* x-directed, TF_SF.cpp:396-640 */
        T->face_start[fs++] = T->npts;
        if (s->pbx) continue;
        int i = e ? T->i2 : T->i1;
        double sh = e ? 0.5 : -0.5;
        for (int j = T->j1; j < T->j2; ++j) for (int k = T->k1; k < T->k2; ++k) tfsf_point(s, T, i, j, k, 3, 0, sh);
        for (int j = T->j1; j < T->j2; ++j) tfsf_point(s, T, i, j, T->k2, 1, 0, sh);
        for (int k = T->k1; k < T->k2; ++k) tfsf_point(s, T, i, T->j2, k, 2, 0, sh);
    }
    for (int e = 0; e < 2; e++) { /* z-directed, TF_SF.cpp:643-887 */
        T->face_start[fs++] = T->npts;
        int k = e ? T->k2 : T->k1;
        double sh = e ? 0.5 : -0.5;
        for (int i = T->i1; i < T->i2; ++i) for (int j = T->j1; j < T->j2; ++j) tfsf_point(s, T, i, j, k, 3, 2, sh);
        if (!s->pby) for (int i = T->i1; i < T->i2; ++i) tfsf_point(s, T, i, T->j2, k, 1, 2, sh);
        if (!s->pbx) for (int j = T->j1; j < T->j2; ++j) tfsf_point(s, T, T->i2, j, k, 2, 2, sh);
    }
    T->face_start[fs] = T->npts;
    /* TF_SF.cpp:889-894 */
    T->CEx = s->dt / (eps0 * s->dx); T->CEy = s->dt / (eps0 * s->dy); T->CEz = s->dt / (eps0 * s->dz);
    T->CHx = s->dt / (mu0 * s->dx); T->CHy = s->dt / (mu0 * s->dy); T->CHz = s->dt / (mu0 * s->dz);
    /* TF_SF.cpp:898-955: waveform tables, 3 spare entries past Nt for late retarded times */
    int Nt = T->Nt;
    if (nt_override > Nt) Nt = nt_override;
    T->Nt = Nt;
    T->Einc = calloc(Nt + 4, sizeof(double));
    T->Hinc = calloc(Nt + 4, sizeof(double));
    double dt = s->dt;
    double nu = C0 / T->lambda, Tp = 1.0 / nu;
    int nT = (int)(Tp / dt);
    double dnu = C0 / (T->lambda - T->dlambda / 2.0) - C0 / (T->lambda + T->dlambda / 2.0);
    double domega = 2.0 * pi * dnu;
    double sig = domega * domega / (16 * log(2));
    int nd = 5 * nT;
    for (int n = 0; n <= Nt + 2; ++n) {
        if (T->cwmode) {
            if (n < nd) {
                T->Einc[n] = exp(-((n * dt) - 5.0 * Tp) * ((n * dt) - 5 * Tp) / (4.0 * Tp * Tp)) * sin(2.0 * pi * nu * (n * dt));
                T->Hinc[n] = exp(-((n * dt + dt / 2.0) - 5.0 * Tp) * ((n * dt + dt / 2.0) - 5 * Tp) / (4.0 * Tp * Tp)) * sin(2.0 * pi * nu * (n * dt + dt / 2.0));
            } else {
                T->Einc[n] = sin(2.0 * pi * nu * (n * dt));
                T->Hinc[n] = sin(2.0 * pi * nu * (n * dt + dt / 2.0));
            }
        } else {
            double t0 = 5 * Tp;
            T->Einc[n] = exp(-sig * ((double)n * dt - t0) * ((double)n * dt - t0)) * sin(2.0 * pi * nu * ((double)n * dt - t0));
            T->Hinc[n] = exp(-sig * ((double)n * dt + dt / 2.0 - t0) * ((double)n * dt + dt / 2.0 - t0)) * sin(2.0 * pi * nu * ((double)n * dt + dt / 2.0 - t0));
        }
    }
    T->ok = 1;
    return 0;
fail:
    return -1;
}

/* TF_SF.cpp:966-986 */
static double interp(const double *tab, double retarded)
{
    if (retarded <= 0) return 0.0;
    int nt = (int)floor(retarded);
    double resid = retarded - (double)nt;
    return ((1.0 - resid) * tab[nt] + resid * tab[nt + 1]);
}

/* TF/SF correction tables by face axis (0 x, 1 y, 2 z).  comp1/comp2 are the
 * tangential components the face's main loop corrects (edge loops correct
 * one of them); the sign is the low face's, the high face flips it.
 *   E: TF_SF.cpp:991-1087   H: TF_SF.cpp:1204-1300 */
static void tfsf_face_apply(orc_sim *s, int n, int is_h, int a, int e, int variant)
{
    tfsf_t *T = &s->tf;
    double *Ex = s->F[0], *Ey = s->F[1], *Ez = s->F[2], *Hx = s->F[3], *Hy = s->F[4], *Hz = s->F[5];
    double *tE[3][2] = {{Ey, Ez}, {Ex, Ez}, {Ex, Ey}};
    double cE[3][2] = {{T->CEx * T->H0z, T->CEx * T->H0y}, {T->CEy * T->H0z, T->CEy * T->H0x}, {T->CEz * T->H0y, T->CEz * T->H0x}};
    double sE[3][2] = {{+1, -1}, {-1, +1}, {+1, -1}};
    double *tH[3][2] = {{Hz, Hy}, {Hz, Hx}, {Hy, Hx}};
    double cH[3][2] = {{T->CHx * T->E0y, T->CHx * T->E0z}, {T->CHy * T->E0x, T->CHy * T->E0z}, {T->CHz * T->E0x, T->CHz * T->E0y}};
    double sH[3][2] = {{+1, -1}, {-1, +1}, {+1, -1}};
    int far = (s->flags & ORC_FIX_TFSF_FARFACE) ? 0 : 1;
    int oa = a == 1 ? 0 : a == 0 ? 1 : 2; /* records: y faces, then x, then z */
    {
        {
            int f = oa * 2 + e;
            double flip = e ? -1.0 : 1.0;
            for (int q = T->face_start[f]; q < T->face_start[f + 1]; q++) {
                const tfsf_pt *pt = &T->pt[q];
                int i = pt->i, j = pt->j, k = pt->k;
                double *c2;
                double (*sg)[2];
                double *(*tg)[2];
                if (is_h) {
                    int off = e ? far : -1;
                    if (a == 0) i += off; else if (a == 1) j += off; else k += off;
                    tg = tH; c2 = cH[a]; sg = sH;
                } else {
                    tg = tE; c2 = cE[a]; sg = sE;
                }
                size_t p = IDX(s, i, j, k);
                for (int c = 0; c < 2; c++) {
                    if (!(pt->part & (1 << c))) continue;
                    double inc, prof;
                    if (is_h) {
                        inc = interp(T->Einc, (double)n - pt->dH[c]);
                        prof = pt->pH[c];
                        /* zero-heap emulation: the reference's x-face H edge loops read
                         * the unset half of the point record (F16) */
                        if ((s->flags & ORC_REF_ZEROHEAP) && a == 0 && pt->part != 3) prof = 0.0;
                    } else {
                        inc = interp(T->Hinc, (double)n - pt->dE[c]);
                        prof = pt->pE[c];
                    }
                    /* TF_SF.cpp:1126-1143, 1375-1392: the z1 face variants
                     * leave the profile factor out */
                    if (variant && a == 2 && e == 0) prof = 1.0;
                    double *F = tg[a][c];
                    if (sg[a][c] * flip < 0) F[p] = F[p] - c2[c] * inc * prof;
                    else F[p] = F[p] + c2[c] * inc * prof;
                }
            }
        }
    }
}

/* the whole update: y faces, then x, then z (TF_SF.cpp:988-1089, 1201-1301) */
static void tfsf_apply(orc_sim *s, int n, int is_h)
{
    const int order[3] = {1, 0, 2};
    for (int oa = 0; oa < 3; oa++)
        for (int e = 0; e < 2; e++) tfsf_face_apply(s, n, is_h, order[oa], e, 0);
}

/* TF_SF::{x,y,z}{1,2}Update{E,H}Field in the order orc_set_tfsf_faces gave */
static void tfsf_apply_faces(orc_sim *s, int n, int is_h)
{
    for (int q = 0; q < s->nfaces; q++) tfsf_face_apply(s, n, is_h, s->faces[q] / 2, s->faces[q] % 2, 1);
}

int orc_set_tfsf_faces(orc_sim *s, const int *faces, int n)
{
    if (n < 0 || n > 64) { snprintf(g_err, sizeof g_err, "0..64 faces"); return -1; }
    for (int q = 0; q < n; q++)
        if (faces[q] < 0 || faces[q] > 5) { snprintf(g_err, sizeof g_err, "face %d not in 0..5", faces[q]); return -1; }
    for (int q = 0; q < n; q++) s->faces[q] = faces[q];
    s->nfaces = n;
    return 0;
}

/* ------------------------------------------------------------------------ */
/* Drude-Lorentz (Drude_Lorentz.cpp:30-148, 3-D file layout, F9/F10 fixed)   */
/* ------------------------------------------------------------------------ */
static int load_dl(orc_sim *s, const char *path)
{
    dl_t *D = &s->dl;
    lines_t L;
    const char *p;
    if (lines_read(path, &L)) FAIL("cannot read DL file %s", path);
    p = lines_next(&L); if (!tok_i(&p, &D->nD) || D->nD < 0 || D->nD > 8) { lines_free(&L); FAIL("dl: nD"); }
    p = lines_next(&L); if (!tok_i(&p, &D->nL) || D->nL < 0 || D->nL > 8) { lines_free(&L); FAIL("dl: nL"); }
    p = lines_next(&L); if (!tok_d(&p, &D->epsInf)) { lines_free(&L); FAIL("dl: epsInf"); }
    D->sigma = 0.0;
    int nD = D->nD, nL = D->nL;
    D->wD = calloc(nD + 1, 8); D->gD = calloc(nD + 1, 8); D->kap = calloc(nD + 1, 8); D->beta = calloc(nD + 1, 8); D->C4 = calloc(nD + 1, 8);
    D->wL = calloc(nL + 1, 8); D->dL = calloc(nL + 1, 8); D->deL = calloc(nL + 1, 8);
    D->alpha = calloc(nL + 1, 8); D->xi = calloc(nL + 1, 8); D->gamL = calloc(nL + 1, 8); D->C5 = calloc(nL + 1, 8); D->C6 = calloc(nL + 1, 8);
    double dt = s->dt, t1 = 0, t2 = 0;
    for (int q = 0; q < nD; q++) {
        p = lines_next(&L);
        if (!tok_d(&p, &D->wD[q]) || !tok_d(&p, &D->gD[q])) { lines_free(&L); FAIL("dl: Drude pole %d", q); }
        D->kap[q] = (1.0 - D->gD[q] * dt / 2.0) / (1.0 + D->gD[q] * dt / 2.0);
        D->beta[q] = (D->wD[q] * D->wD[q] * eps0 * dt / 2.0) / (1.0 + D->gD[q] * dt / 2.0);
        t1 = t1 + D->beta[q];
    }
    for (int q = 0; q < nL; q++) {
        p = lines_next(&L);
        if (!tok_d(&p, &D->wL[q]) || !tok_d(&p, &D->dL[q]) || !tok_d(&p, &D->deL[q])) { lines_free(&L); FAIL("dl: Lorentz pole %d", q); }
        D->alpha[q] = (2.0 - D->wL[q] * D->wL[q] * dt * dt) / (1.0 + D->dL[q] * dt);
        D->xi[q] = (D->dL[q] * dt - 1.0) / (1.0 + D->dL[q] * dt);
        D->gamL[q] = (eps0 * D->deL[q] * D->wL[q] * D->wL[q] * dt * dt) / (1.0 + D->dL[q] * dt);
        t2 = t2 + D->gamL[q];
    }
    lines_free(&L);
    double den = 2.0 * eps0 * D->epsInf + D->sigma * dt + dt * t1 + t2 / 2.0;
    D->C1 = (2.0 * eps0 * D->epsInf - D->sigma * dt - dt * t1) / den;
    D->C2 = t2 / 2.0 / den;
    D->C3 = 2.0 * dt / den;
    for (int q = 0; q < nD; q++) D->C4[q] = (1.0 + D->kap[q]) * dt / den;
    for (int q = 0; q < nL; q++) { D->C5[q] = (1.0 + D->alpha[q]) * dt / den; D->C6[q] = D->xi[q] * dt / den; }
    /* cell list: ID==2 cells in the E-update range, i-j-k order */
    int ie0 = s->pbx ? 0 : 1, je0 = s->pby ? 0 : 1, cnt = 0;
    for (int i = ie0; i < s->Nx; i++) for (int j = je0; j < s->Ny; j++) for (int k = 1; k < s->Nz; k++)
        if (s->ID[IDX(s, i, j, k)] == 2) cnt++;
    D->ncell = cnt;
    D->cell = calloc(cnt + 1, sizeof(int));
    D->st = calloc((size_t)(cnt + 1) * 3 * (nD + 2 * nL + 1), sizeof(double));
    cnt = 0;
    for (int i = ie0; i < s->Nx; i++) for (int j = je0; j < s->Ny; j++) for (int k = 1; k < s->Nz; k++)
        if (s->ID[IDX(s, i, j, k)] == 2) D->cell[cnt++] = (int)IDX(s, i, j, k);
    D->ok = 1;
    return 0;
fail:
    return -1;
}

/* ------------------------------------------------------------------------ */
/* Four-level gain: 3-D single-electron model (Chang & Taflove, Opt. Express  */
/* 12, 3827 (2004)) with a pump rate, on ID==5 cells.  File layout follows   */
/* FourLevelLaser.cpp:37-93 plus a 15th line holding the pump rate W_p.      */
/* ------------------------------------------------------------------------ */
static int load_gain(orc_sim *s, const char *path)
{
    gain_t *G = &s->gn;
    lines_t L;
    const char *p;
    double v[14];
    if (lines_read(path, &L)) FAIL("cannot read gain file %s", path);
    for (int q = 0; q < 15; q++) {
        p = lines_next(&L);
        if (q < 4) continue; /* distribution, NumSources, i1 i2, j1 j2: ID==5 cells are the sites */
        if (!tok_d(&p, q < 14 ? &v[q] : &G->Wp)) { lines_free(&L); FAIL("gain: line %d", q + 1); }
    }
    lines_free(&L);
    G->lambda_a = v[5]; G->gamma_a = v[6]; G->t32 = v[9]; G->t21 = v[10]; G->t10 = v[11]; G->Ntot = v[12];
    if (G->t32 <= 0 || G->t21 <= 0 || G->t10 <= 0 || G->lambda_a <= 0) FAIL("gain: lifetimes/wavelength must be positive");
    double dt = s->dt;
    G->omega_a = 2.0 * pi * C0 / G->lambda_a;
    G->kappa_a = 6.0 * pi * eps0 * C0 * C0 * C0 / (G->omega_a * G->omega_a * G->t21);
    /* FourLevelLaser.cpp:100-102 */
    G->CPa1 = (2.0 - G->omega_a * G->omega_a * dt * dt) / (1.0 + G->gamma_a * dt / 2.0);
    G->CPa2 = (G->gamma_a * dt / 2.0 - 1.0) / (1.0 + G->gamma_a * dt / 2.0);
    G->CPa3 = dt * dt * G->kappa_a / (1.0 + G->gamma_a * dt / 2.0);
    int ie0 = s->pbx ? 0 : 1, je0 = s->pby ? 0 : 1, cnt = 0;
    for (int i = ie0; i < s->Nx; i++) for (int j = je0; j < s->Ny; j++) for (int k = 1; k < s->Nz; k++)
        if (s->ID[IDX(s, i, j, k)] == 5) cnt++;
    G->ncell = cnt;
    G->cell = calloc(cnt + 1, sizeof(int));
    G->P = calloc((size_t)(cnt + 1) * 3, 8); G->P1 = calloc((size_t)(cnt + 1) * 3, 8);
    G->Eprev = calloc((size_t)(cnt + 1) * 3, 8); G->N = calloc((size_t)(cnt + 1) * 4, 8);
    cnt = 0;
    for (int i = ie0; i < s->Nx; i++) for (int j = je0; j < s->Ny; j++) for (int k = 1; k < s->Nz; k++)
        if (s->ID[IDX(s, i, j, k)] == 5) { G->N[(size_t)cnt * 4] = G->Ntot; G->cell[cnt++] = (int)IDX(s, i, j, k); }
    G->ok = 1;
    return 0;
fail:
    return -1;
}

/* ------------------------------------------------------------------------ */
/* Sources (Sources.cpp:29-76)                                               */
/* ------------------------------------------------------------------------ */
static int load_sources(orc_sim *s, const orc_config *cfg)
{
    s->nsrc = cfg->n_src;
    if (!s->nsrc) return 0;
    if (!cfg->source) FAIL("point sources need a source file");
    lines_t L;
    const char *p;
    double lambda, fwhm, delay;
    if (lines_read(cfg->source, &L)) FAIL("cannot read source file %s", cfg->source);
    p = lines_next(&L); if (!tok_d(&p, &lambda)) { lines_free(&L); FAIL("source: lambda"); }
    p = lines_next(&L); if (!tok_d(&p, &fwhm)) { lines_free(&L); FAIL("source: FWHM"); }
    p = lines_next(&L); if (!tok_d(&p, &delay)) { lines_free(&L); FAIL("source: delay"); }
    lines_free(&L);
    s->src = malloc(sizeof(int) * 4 * s->nsrc);
    s->swave = calloc((size_t)s->nsrc * (s->nt + 1), sizeof(double));
    double dt = s->dt, nu = C0 / lambda, Tp = 1.0 / nu;
    int nT = (int)(Tp / dt), nw = 2 * nT, nd = 4 * nw;
    double pinudt = 2.0 * pi * nu * dt;
    for (int q = 0; q < s->nsrc; q++) {
        int i = cfg->src_ijkc[4 * q], j = cfg->src_ijkc[4 * q + 1], k = cfg->src_ijkc[4 * q + 2], c = cfg->src_ijkc[4 * q + 3];
        if (i < 0 || j < 0 || k < 0 || i > s->Nx || j > s->Ny || k > s->Nz || c < 0 || c > 2) FAIL("source %d out of range", q);
        memcpy(s->src + 4 * q, cfg->src_ijkc + 4 * q, 4 * sizeof(int));
        double A = cfg->src_amp ? cfg->src_amp[q] : 1.0, ph = cfg->src_phase ? cfg->src_phase[q] : 0.0;
        for (int n = 0; n <= s->nt; n++) {
            double w;
            if (s->flags & ORC_SRC_REF_PULSE) {
                double retard = n - 0.0 / C0 / dt; /* Sources.cpp:72-76, location 0 */
                w = (A * exp(-(double)(retard - nd) * (retard - nd) / (nw * nw)) * sin(pinudt * (retard - nd)));
            } else {
                double t = (double)n * dt - delay;
                w = A * exp(-4.0 * log(2.0) * t * t / (fwhm * fwhm)) * sin(2.0 * pi * nu * t + ph);
            }
            s->swave[(size_t)q * (s->nt + 1) + n] = w;
        }
    }
    return 0;
fail:
    return -1;
}

/* ------------------------------------------------------------------------ */
/* Ports (Ports.cpp:28-86 sphere, :201-306 plane)                            */
/* ------------------------------------------------------------------------ */
static int load_port(orc_sim *s, port_t *P, const char *path, int kind)
{
    lines_t L;
    const char *p;
    P->kind = kind;
    if (lines_read(path, &L)) FAIL("cannot read port file %s", path);
    double dx = s->dx, dy = s->dy, dz = s->dz;
    if (kind == ORC_PORT_SPHERE) {
        double r, xc, yc, zc, t1, t2, p1, p2;
        p = lines_next(&L); if (!tok_d(&p, &r)) { lines_free(&L); FAIL("sphere port: r"); }
        p = lines_next(&L); if (!tok_d(&p, &xc) || !tok_d(&p, &yc) || !tok_d(&p, &zc)) { lines_free(&L); FAIL("sphere port: centre"); }
        p = lines_next(&L); if (!tok_d(&p, &t1) || !tok_d(&p, &t2)) { lines_free(&L); FAIL("sphere port: theta"); }
        p = lines_next(&L); if (!tok_d(&p, &p1) || !tok_d(&p, &p2)) { lines_free(&L); FAIL("sphere port: phi"); }
        lines_free(&L);
        t1 = t1 * pi / 180.0; t2 = t2 * pi / 180.0; p1 = p1 * pi / 180.0; p2 = p2 * pi / 180.0;
        double dphi = dx / r, dtheta = dz / r;
        int n = 0;
        for (double phi = p1; phi <= p2; phi = phi + dphi)
            for (double th = t1; th <= t2; th = th + dtheta) n++;
        P->npts = n;
        P->ijk = malloc(sizeof(int) * 3 * (n + 1));
        n = 0;
        for (double phi = p1; phi <= p2; phi = phi + dphi)
            for (double th = t1; th <= t2; th = th + dtheta) {
                P->ijk[3 * n] = (int)((xc + r * sin(th) * cos(phi)) / dx);
                P->ijk[3 * n + 1] = (int)((yc + r * sin(th) * sin(phi)) / dy);
                P->ijk[3 * n + 2] = (int)((zc + r * cos(th)) / dz);
                n++;
            }
    } else {
        char nd[8] = {0};
        double pos, w1, w2, h1, h2;
        p = lines_next(&L); if (!tok(&p, nd, sizeof nd)) { lines_free(&L); FAIL("plane port: normal"); }
        p = lines_next(&L); if (!tok_d(&p, &pos)) { lines_free(&L); FAIL("plane port: position"); }
        p = lines_next(&L); if (!tok_d(&p, &w1) || !tok_d(&p, &w2)) { lines_free(&L); FAIL("plane port: w"); }
        p = lines_next(&L); if (!tok_d(&p, &h1) || !tok_d(&p, &h2)) { lines_free(&L); FAIL("plane port: h"); }
        lines_free(&L);
        int a1, a2, b1, b2, c1;
        if (nd[0] == 'x') { c1 = (int)(pos / dx); a1 = (int)(w1 / dy); a2 = (int)(w2 / dy); b1 = (int)(h1 / dz); b2 = (int)(h2 / dz); }
        else if (nd[0] == 'y') { c1 = (int)(pos / dy); a1 = (int)(w1 / dx); a2 = (int)(w2 / dx); b1 = (int)(h1 / dz); b2 = (int)(h2 / dz); }
        else if (nd[0] == 'z') { c1 = (int)(pos / dz); a1 = (int)(w1 / dx); a2 = (int)(w2 / dx); b1 = (int)(h1 / dy); b2 = (int)(h2 / dy); }
        else FAIL("plane port: normal must be x, y or z");
        if (a2 < a1 || b2 < b1) FAIL("plane port: empty extent");
        int n = (a2 - a1 + 1) * (b2 - b1 + 1);
        P->npts = n;
        P->ijk = malloc(sizeof(int) * 3 * (n + 1));
        n = 0;
        for (int b = b1; b <= b2; b++)
            for (int a = a1; a <= a2; a++) {
                int *q = P->ijk + 3 * n++;
                if (nd[0] == 'x') { q[0] = c1; q[1] = a; q[2] = b; }
                else if (nd[0] == 'y') { q[0] = a; q[1] = c1; q[2] = b; }
                else { q[0] = a; q[1] = b; q[2] = c1; }
            }
    }
    for (int q = 0; q < P->npts; q++) {
        int *ijk = P->ijk + 3 * q;
        if (s->pbx && ijk[0] == s->Nx) ijk[0] = 0;
        if (s->pby && ijk[1] == s->Ny) ijk[1] = 0;
        if (ijk[0] < 0 || ijk[1] < 0 || ijk[2] < 0 || ijk[0] > s->Nx || ijk[1] > s->Ny || ijk[2] > s->Nz)
            FAIL("port %s: point %d outside the grid", path, q);
    }
    if (s->record) P->samples = calloc((size_t)s->nt * P->npts * 6, sizeof(double));
    if (s->nf) P->dft = calloc((size_t)P->npts * s->nf * 12, sizeof(double));
    P->inten = calloc((size_t)s->nt + 1, sizeof(double));
    return 0;
fail:
    return -1;
}

/* Ports.cpp:342-353 gather, then the running DFT and intensity sum */
static void ports_sample(orc_sim *s, int n)
{
    for (int q = 0; q < s->nports; q++) {
        port_t *P = &s->port[q];
        double isum = 0.0;
        for (int t = 0; t < P->npts; t++) {
            size_t p = IDX(s, P->ijk[3 * t], P->ijk[3 * t + 1], P->ijk[3 * t + 2]);
            double v[6];
            for (int c = 0; c < 6; c++) v[c] = s->F[c][p];
            if (P->samples) memcpy(P->samples + ((size_t)n * P->npts + t) * 6, v, sizeof v);
            isum = isum + (v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
            if (P->dft && n % s->stride == 0) {
                for (int f = 0; f < s->nf; f++) {
                    const double *wE = s->twE + ((size_t)f * s->nt + n) * 2, *wH = s->twH + ((size_t)f * s->nt + n) * 2;
                    double *acc = P->dft + ((size_t)t * s->nf + f) * 12;
                    for (int c = 0; c < 6; c++) {
                        const double *w = c < 3 ? wE : wH;
                        acc[2 * c] = acc[2 * c] + v[c] * w[0];
                        acc[2 * c + 1] = acc[2 * c + 1] - v[c] * w[1];
                    }
                }
            }
        }
        P->inten[n] = isum;
    }
}

/* ------------------------------------------------------------------------ */
/* create / step / accessors                                                 */
/* ------------------------------------------------------------------------ */
orc_sim *orc_create(const orc_config *cfg)
{
    orc_sim *s = calloc(1, sizeof(orc_sim));
    g_err[0] = 0;
    s->flags = cfg->flags;
    s->periodic = cfg->periodic;
    s->pbx = cfg->periodic & 1;
    s->pby = (cfg->periodic >> 1) & 1;
    if (load_structure(s, cfg->material, cfg->structure)) goto fail;
    init_coefficients(s);
    for (int c = 0; c < 6; c++) s->F[c] = calloc(s->ncell, sizeof(double));
    if (cfg->pml && load_cpml(s, cfg->pml)) goto fail;
    s->nt = cfg->nt;
    if (cfg->tfsf) {
        if (load_tfsf(s, cfg->tfsf, cfg->nt)) goto fail;
        if (s->nt <= 0) s->nt = s->tf.Nt;
    }
    if (s->nt <= 0) FAIL("number of steps not given");
    if (cfg->dl && load_dl(s, cfg->dl)) goto fail;
    if (cfg->gain && load_gain(s, cfg->gain)) goto fail;
    if (load_sources(s, cfg)) goto fail;
    s->nf = cfg->n_freq;
    s->stride = cfg->dft_stride > 0 ? cfg->dft_stride : 1;
    s->record = cfg->record_ports;
    if (s->nf) {
        s->freqs = malloc(sizeof(double) * s->nf);
        memcpy(s->freqs, cfg->freqs, sizeof(double) * s->nf);
        s->twE = malloc(sizeof(double) * 2 * s->nf * s->nt);
        s->twH = malloc(sizeof(double) * 2 * s->nf * s->nt);
        for (int f = 0; f < s->nf; f++)
            for (int n = 0; n < s->nt; n++) {
                double tE = ((double)n + 1.0) * s->dt, tH = ((double)n + 0.5) * s->dt;
                double aE = 2.0 * pi * s->freqs[f] * tE, aH = 2.0 * pi * s->freqs[f] * tH;
                s->twE[((size_t)f * s->nt + n) * 2] = cos(aE); s->twE[((size_t)f * s->nt + n) * 2 + 1] = sin(aE);
                s->twH[((size_t)f * s->nt + n) * 2] = cos(aH); s->twH[((size_t)f * s->nt + n) * 2 + 1] = sin(aH);
            }
    }
    s->nports = cfg->n_ports;
    s->port = calloc(s->nports + 1, sizeof(port_t));
    for (int q = 0; q < s->nports; q++)
        if (load_port(s, &s->port[q], cfg->port_files[q], cfg->port_kinds ? cfg->port_kinds[q] : 0)) goto fail;
    return s;
fail:
    orc_destroy(s);
    return NULL;
}

void orc_destroy(orc_sim *s)
{
    if (!s) return;
    free(s->epsR); free(s->muR); free(s->sig); free(s->sigM);
    free(s->aE); free(s->bE); free(s->aH); free(s->bH);
    for (int t = 0; t < 12; t++) free(s->cst[t]);
    free(s->ID);
    for (int c = 0; c < 6; c++) free(s->F[c]);
    for (int a = 0; a < 3; a++) for (int e = 0; e < 2; e++) {
        free(s->pml.bE[a][e]); free(s->pml.cE[a][e]); free(s->pml.bH[a][e]); free(s->pml.cH[a][e]);
        for (int q = 0; q < 2; q++) { free(s->pml.qE[a][e][q]); free(s->pml.qH[a][e][q]); }
    }
    free(s->tf.pt); free(s->tf.Einc); free(s->tf.Hinc);
    dl_t *D = &s->dl;
    free(D->wD); free(D->gD); free(D->kap); free(D->beta); free(D->C4); free(D->wL); free(D->dL); free(D->deL);
    free(D->alpha); free(D->xi); free(D->gamL); free(D->C5); free(D->C6); free(D->cell); free(D->st);
    gain_t *G = &s->gn;
    free(G->cell); free(G->P); free(G->P1); free(G->Eprev); free(G->N);
    free(s->src); free(s->swave);
    for (int q = 0; q < s->nports && s->port; q++) {
        free(s->port[q].ijk); free(s->port[q].samples); free(s->port[q].dft); free(s->port[q].inten);
    }
    free(s->port); free(s->freqs); free(s->twE); free(s->twH);
    free(s);
}

/* main.cpp:267-343 step order */
int orc_step(orc_sim *s, int count)
{
    for (int c = 0; c < count; c++) {
        int n = s->ndone;
        if (n >= s->nt) { snprintf(g_err, sizeof g_err, "run finished (%d steps)", s->nt); return -1; }
        basic_H(s);
        if (s->tf.ok) { if (s->nfaces) tfsf_apply_faces(s, n, 1); else tfsf_apply(s, n, 1); }
        if (s->pml.ok) cpml_update_H(s);
        basic_E(s);
        if (s->tf.ok) { if (s->nfaces) tfsf_apply_faces(s, n, 0); else tfsf_apply(s, n, 0); }
        for (int q = 0; q < s->nsrc; q++) {
            size_t p = IDX(s, s->src[4 * q], s->src[4 * q + 1], s->src[4 * q + 2]);
            double *Fc = s->F[s->src[4 * q + 3]];
            Fc[p] = Fc[p] + s->swave[(size_t)q * (s->nt + 1) + n];
        }
        if (s->pml.ok) cpml_update_E(s);
        ports_sample(s, n);
        s->ndone++;
    }
    return 0;
}

int orc_steps_done(const orc_sim *s) { return s->ndone; }

void orc_dims(const orc_sim *s, int *nxyz, double *dxyzt)
{
    nxyz[0] = s->Nx; nxyz[1] = s->Ny; nxyz[2] = s->Nz; nxyz[3] = s->nt;
    dxyzt[0] = s->dx; dxyzt[1] = s->dy; dxyzt[2] = s->dz; dxyzt[3] = s->dt;
}
int orc_nmedia(const orc_sim *s) { return s->NoMedia; }
void orc_get_id(const orc_sim *s, short *out) { memcpy(out, s->ID, s->ncell * sizeof(short)); }
void orc_get_coef(const orc_sim *s, int which, double *out)
{
    const double *src[4] = {s->aE, s->bE, s->aH, s->bH};
    memcpy(out, src[which & 3], sizeof(double) * s->NoMedia);
}
int orc_get_stretch(const orc_sim *s, int which, double *out)
{
    if (which < 0 || which > 11) return 0;
    if (out) memcpy(out, s->cst[which], sizeof(double) * s->cstlen[which]);
    return s->cstlen[which];
}
int orc_get_cpml(const orc_sim *s, int which, double *b, double *c)
{
    int a = which / 4, e = (which / 2) & 1, h = which & 1;
    if (!s->pml.ok || a > 2) return 0;
    int len = h ? s->pml.lenH[a][e] : s->pml.lenE[a][e];
    const double *bb = h ? s->pml.bH[a][e] : s->pml.bE[a][e], *cc = h ? s->pml.cH[a][e] : s->pml.cE[a][e];
    if (b && len) memcpy(b, bb, sizeof(double) * len);
    if (c && len) memcpy(c, cc, sizeof(double) * len);
    return len;
}
int orc_get_tfsf_points(const orc_sim *s, int *ijk, double *tab)
{
    if (!s->tf.ok) return 0;
    for (int q = 0; q < s->tf.npts && (ijk || tab); q++) {
        const tfsf_pt *p = &s->tf.pt[q];
        if (ijk) { ijk[4 * q] = p->i; ijk[4 * q + 1] = p->j; ijk[4 * q + 2] = p->k; ijk[4 * q + 3] = p->part; }
        if (tab)
            for (int c = 0; c < 2; c++) {
                double *t = tab + 8 * q + 4 * c;
                t[0] = p->dE[c]; t[1] = p->pE[c]; t[2] = p->dH[c]; t[3] = p->pH[c];
            }
    }
    return s->tf.npts;
}
int orc_get_tfsf_wave(const orc_sim *s, double *einc, double *hinc)
{
    if (!s->tf.ok) return 0;
    if (einc) memcpy(einc, s->tf.Einc, sizeof(double) * (s->tf.Nt + 1));
    if (hinc) memcpy(hinc, s->tf.Hinc, sizeof(double) * (s->tf.Nt + 1));
    return s->tf.Nt + 1;
}
void orc_get_tfsf_vec(const orc_sim *s, double *o)
{
    const tfsf_t *T = &s->tf;
    double v[12] = {T->E0x, T->E0y, T->E0z, T->H0x, T->H0y, T->H0z, T->CEx, T->CEy, T->CEz, T->CHx, T->CHy, T->CHz};
    memcpy(o, v, sizeof v);
}
int orc_port_npts(const orc_sim *s, int port) { return (port >= 0 && port < s->nports) ? s->port[port].npts : 0; }
void orc_port_points(const orc_sim *s, int port, int *ijk) { memcpy(ijk, s->port[port].ijk, sizeof(int) * 3 * s->port[port].npts); }
void orc_port_samples(const orc_sim *s, int port, double *out)
{
    const port_t *P = &s->port[port];
    if (P->samples) memcpy(out, P->samples, sizeof(double) * (size_t)s->ndone * P->npts * 6);
}
void orc_port_dft(const orc_sim *s, int port, double *out)
{
    const port_t *P = &s->port[port];
    if (P->dft) memcpy(out, P->dft, sizeof(double) * (size_t)P->npts * s->nf * 12);
}
void orc_port_intensity(const orc_sim *s, int port, double *out)
{
    memcpy(out, s->port[port].inten, sizeof(double) * (size_t)s->ndone);
}
void orc_get_field(const orc_sim *s, int comp, double *out) { memcpy(out, s->F[comp], s->ncell * sizeof(double)); }
void orc_set_field(orc_sim *s, int comp, const double *in) { memcpy(s->F[comp], in, s->ncell * sizeof(double)); }
int orc_gain_ncells(const orc_sim *s) { return s->gn.ncell; }
void orc_gain_pop(const orc_sim *s, double *out) { memcpy(out, s->gn.N, sizeof(double) * 4 * (size_t)s->gn.ncell); }
int orc_dl_ncells(const orc_sim *s) { return s->dl.ncell; }
int orc_dl_state(const orc_sim *s, double *out)
{
    const dl_t *D = &s->dl;
    if (out && D->ok) memcpy(out, D->st, sizeof(double) * (size_t)D->ncell * 3 * (D->nD + 2 * D->nL + 1));
    return D->ok ? D->ncell : 0;
}
void orc_dl_poles(const orc_sim *s, int *nD_nL)
{
    nD_nL[0] = s->dl.ok ? s->dl.nD : 0;
    nD_nL[1] = s->dl.ok ? s->dl.nL : 0;
}
int orc_gain_state(const orc_sim *s, double *out)
{
    const gain_t *G = &s->gn;
    if (out && G->ok)
        for (int c = 0; c < G->ncell; c++)
            for (int q = 0; q < 3; q++) {
                out[(size_t)c * 6 + q] = G->P[(size_t)c * 3 + q];
                out[(size_t)c * 6 + 3 + q] = G->P1[(size_t)c * 3 + q];
            }
    return G->ok ? G->ncell : 0;
}
int orc_special_cells(const orc_sim *s, int kind, int *ijk)
{
    int n = kind ? (s->gn.ok ? s->gn.ncell : 0) : (s->dl.ok ? s->dl.ncell : 0);
    const int *cell = kind ? s->gn.cell : s->dl.cell;
    for (int r = 0; ijk && r < n; r++) {
        size_t p = (size_t)cell[r];
        ijk[3 * r] = (int)(p / ((size_t)(s->Ny + 1) * (s->Nz + 1)));
        ijk[3 * r + 1] = (int)((p / (size_t)(s->Nz + 1)) % (size_t)(s->Ny + 1));
        ijk[3 * r + 2] = (int)(p % (size_t)(s->Nz + 1));
    }
    return n;
}

int orc_source_wave(const orc_sim *s, int q, double *out)
{
    if (q < 0 || q >= s->nsrc) return 0;
    if (out) memcpy(out, s->swave + (size_t)q * (s->nt + 1), sizeof(double) * (s->nt + 1));
    return s->nt + 1;
}
int orc_dl_coefs(const orc_sim *s, double *o)
{
    const dl_t *D = &s->dl;
    if (!D->ok) return 0;
    int n = 0;
    o[n++] = D->C1; o[n++] = D->C2; o[n++] = D->C3;
    for (int q = 0; q < D->nD; q++) o[n++] = D->C4[q];
    for (int q = 0; q < D->nL; q++) o[n++] = D->C5[q];
    for (int q = 0; q < D->nL; q++) o[n++] = D->C6[q];
    for (int q = 0; q < D->nD; q++) o[n++] = D->kap[q];
    for (int q = 0; q < D->nD; q++) o[n++] = D->beta[q];
    for (int q = 0; q < D->nL; q++) o[n++] = D->alpha[q];
    for (int q = 0; q < D->nL; q++) o[n++] = D->xi[q];
    for (int q = 0; q < D->nL; q++) o[n++] = D->gamL[q];
    return n;
}

/* single-cell DL recurrence: E^{n+1} = C1 E^n + C2 E^{n-1} + C3 curl^{n+1/2} - sum C4 JD - sum(C5 JL + C6 JL1) */
int orc_dl_0d(const orc_sim *s, int nsteps, const double *curl, double *e_out)
{
    const dl_t *D = &s->dl;
    if (!D->ok) return -1;
    double JD[8] = {0}, JL[8] = {0}, JL1[8] = {0}, E = 0, E1 = 0;
    for (int n = 0; n < nsteps; n++) {
        double temp = E, t1 = 0, t2 = 0;
        for (int q = 0; q < D->nD; q++) t1 = t1 + D->C4[q] * JD[q];
        for (int q = 0; q < D->nL; q++) t2 = t2 + D->C5[q] * JL[q] + D->C6[q] * JL1[q];
        E = D->C1 * E + D->C2 * E1 + D->C3 * curl[n] - t1 - t2;
        for (int q = 0; q < D->nD; q++) JD[q] = D->kap[q] * JD[q] + D->beta[q] * (E + temp);
        for (int q = 0; q < D->nL; q++) {
            double t0 = JL[q];
            JL[q] = D->alpha[q] * JL[q] + D->xi[q] * JL1[q] + D->gamL[q] * (E - E1) / 2.0 / s->dt;
            JL1[q] = t0;
        }
        E1 = temp;
        e_out[n] = E;
    }
    return 0;
}

/* single gain cell with vacuum background: E^{n+1} = efield[n] - (bE0/dt) dP */
int orc_gain_0d(const orc_sim *s, int nsteps, const double *efield, double *pop_out, double *p_out)
{
    const gain_t *G = &s->gn;
    if (!G->ok) return -1;
    double N[4] = {G->Ntot, 0, 0, 0}, P = 0, P1 = 0, E = 0, dt = s->dt;
    double coef = s->bE[0] / dt;
    for (int n = 0; n < nsteps; n++) {
        double Pn = P;
        P = G->CPa1 * Pn + G->CPa2 * P1 + G->CPa3 * (N[1] - N[2]) * E;
        P1 = Pn;
        double dP = P - P1;
        double En = efield[n] - coef * dP;
        double R = ((En + E) * dP) / (2.0 * dt) / (hPlanck / (2.0 * pi) * G->omega_a);
        double h32 = dt / (2.0 * G->t32), h21 = dt / (2.0 * G->t21), h10 = dt / (2.0 * G->t10);
        double N3 = (N[3] * (1.0 - h32) + dt * G->Wp * N[0]) / (1.0 + h32);
        double N2 = (N[2] * (1.0 - h21) + h32 * (N[3] + N3) + dt * R) / (1.0 + h21);
        double N1 = (N[1] * (1.0 - h10) + h21 * (N[2] + N2) - dt * R) / (1.0 + h10);
        N[3] = N3; N[2] = N2; N[1] = N1; N[0] = G->Ntot - N1 - N2 - N3;
        E = En;
        if (pop_out) memcpy(pop_out + 4 * n, N, sizeof N);
        if (p_out) p_out[n] = P;
    }
    return 0;
}
