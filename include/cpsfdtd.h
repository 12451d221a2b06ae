/*
 * cpsfdtd.h -- C ABI of the MI355X FDTD engine (libcpsfdtd.so).
 *
 * One opaque handle per simulation; its state (fields, CPML psi, dispersive
 * and gain auxiliaries, port accumulators) stays resident in HBM.  Every call
 * on a handle is ordered on the handle's HIP stream; calls that return host
 * data synchronise that stream.  Return value 0 = success, negative = error,
 * cfd_last_error() then holds a message (thread-local).
 *
 * Field arrays crossing this boundary are FP64 in the reference's order:
 * (Nx+1)(Ny+1)(Nz+1) values, [i][j][k], k fastest (Fields.cpp:99-131).
 *
 * The reference has no FFI; the entry points replace its C++ class calls in
 * main.cpp's time loop, cited per function.
 */
#ifndef CPSFDTD_H
#define CPSFDTD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* behaviour switches (bit values shared with the oracle) */
#define CFD_FIX_CPML_HSIGN   1u  /* CPML.cpp:643-682 H-side y/z psi sign -> Roden-Gedney */
#define CFD_FIX_TFSF_FARFACE 2u  /* TF_SF.cpp:1222,1255,1288 far-face H at index+0 */
#define CFD_SRC_REF_PULSE    4u  /* point sources follow Sources.cpp:72-76 Pulse() */
#define CFD_REF_ZEROHEAP     8u  /* reproduce the reference's reads of zero-filled unset memory */
#define CFD_FIX_TFSF_YEE    16u  /* TF_SF.cpp:155-887 incident values at each component's Yee position */

#define CFD_PORT_PLANE  0
#define CFD_PORT_SPHERE 1

#define CFD_F32 32
#define CFD_F64 64

typedef struct {
    const char *material;   /* Structure(const char*)              Structure.cpp:26   */
    const char *structure;  /* Structure::generate_mesh            Structure.cpp:68   */
    const char *pml;        /* CPML / CFS_PML(Structure&, char*)   CPML.cpp:29        NULL = none */
    const char *tfsf;       /* TF_SF(Structure&, char*)            TF_SF.cpp:29       NULL = none */
    const char *dl;         /* Drude_Lorentz(Structure&, char*)    Drude_Lorentz.cpp:30 (3-D layout) */
    const char *gain;       /* FourLevelLaser(Structure&, char*)   FourLevelLaser.cpp:30 (+ pump line) */
    const char *source;     /* Source(double, char*)               Sources.cpp:29     */
    int n_ports;
    const char *const *port_files;   /* SphericalPort / PlainPort(Structure&, const char*) Ports.cpp:28,201 */
    const int *port_kinds;           /* CFD_PORT_PLANE / CFD_PORT_SPHERE */
    int n_src;                       /* soft point sources (main.cpp:284-288) */
    const int *src_ijkc;             /* i j k comp(0..2) per source */
    const double *src_amp;
    const double *src_phase;
    int nt;                          /* steps; <= 0 takes Nt from TF_SF.txt (main.cpp:267) */
    int n_freq;                      /* running-DFT frequencies at the ports (Hz) */
    const double *freqs;
    int dft_stride;                  /* accumulate every dft_stride-th step */
    int record_ports;                /* keep every step's port samples on the device */
    int periodic;                    /* bit0 x, bit1 y periodic (PBC.h:21-38 stub, built new) */
    unsigned flags;                  /* CFD_* switches */
    int precision;                   /* CFD_F32 or CFD_F64 field storage and update */
} cfd_config;

typedef struct cfd_sim cfd_sim;

const char *cfd_last_error(void);
int cfd_device_info(char *name, int cap, int *n_cu, size_t *hbm_bytes);

/* Structure + Field + CPML::InitializeCPML + TF_SF + Drude_Lorentz +
 * FourLevelLaser + ports construction (main.cpp:115-147), then upload. */
int cfd_create(const cfd_config *cfg, cfd_sim **out);
int cfd_destroy(cfd_sim *s);
/* nxyz_nt: Nx Ny Nz nt; d: dx dy dz dt; counts: n_ports, n_dl_cells, n_gain_cells, n_tfsf_targets */
int cfd_info(cfd_sim *s, int *nxyz_nt, double *d, int *counts);
size_t cfd_device_bytes(const cfd_sim *s);

/* One full time step per count, fused kernels (main.cpp:267-343 order:
 * H, TF/SF-H, CPML-H, E, DL, gain, TF/SF-E, sources, CPML-E, ports). */
int cfd_step(cfd_sim *s, int count);
int cfd_steps_done(cfd_sim *s);
int cfd_sync(cfd_sim *s);
int cfd_query(cfd_sim *s);
/* number of non-finite (NaN, inf) field values now (0: healthy); reads every
 * field once.  No reference counterpart (the reference checks nothing). */
int cfd_check_finite(cfd_sim *s);               /* 1 = all queued work finished, 0 = still running (no wait) */

/* Single-module passes in the reference's granularity, for drivers that keep
 * main.cpp's call sequence.  n is the time index the reference passes. */
int cfd_basic_h(cfd_sim *s);            /* BasicHFieldUpdate      BasicFDTD.cpp:67-81   */
int cfd_basic_e(cfd_sim *s);            /* BasicEFieldUpdate      BasicFDTD.cpp:25-45   */
int cfd_cpml_h(cfd_sim *s);             /* CPML::UpdateCPMLHFields CPML.cpp:609-689     */
int cfd_cpml_e(cfd_sim *s);             /* CPML::UpdateCPMLEFields CPML.cpp:529-607     */
int cfd_tfsf_h(cfd_sim *s, int n);      /* TF_SF::UpdateHField    TF_SF.cpp:1201-1301   */
int cfd_tfsf_e(cfd_sim *s, int n);      /* TF_SF::UpdateEField    TF_SF.cpp:988-1089    */
/* One face of the TF/SF box alone: TF_SF::x1UpdateHField ... z2UpdateHField
 * (TF_SF.cpp:1302-1410) and x1UpdateEField ... z2UpdateEField
 * (TF_SF.cpp:1090-1199).  The z1 forms leave out the transverse profile, as
 * the reference's do. */
#define CFD_FACE_X1 0
#define CFD_FACE_X2 1
#define CFD_FACE_Y1 2
#define CFD_FACE_Y2 3
#define CFD_FACE_Z1 4
#define CFD_FACE_Z2 5
int cfd_tfsf_face_h(cfd_sim *s, int face, int n);
int cfd_tfsf_face_e(cfd_sim *s, int face, int n);
int cfd_dl_e(cfd_sim *s);               /* Drude_Lorentz::EFieldUpdate Drude_Lorentz.cpp:153-227 (3-D) */
int cfd_gain_pol(cfd_sim *s);           /* FourLevelLaser::UpdatePolarizationDensities FourLevelLaser.cpp:270-298 */
int cfd_gain_e(cfd_sim *s);             /* FourLevelLaser::UpdateElectricField FourLevelLaser.cpp:300-308 */
int cfd_gain_pop(cfd_sim *s);           /* FourLevelLaser::UpdateStatesPopulations FourLevelLaser.cpp:310-326 */
int cfd_sources(cfd_sim *s, int n);     /* soft point sources     main.cpp:284-288      */
int cfd_ports_sample(cfd_sim *s, int n);/* *Port::Update_Fields_on_Port Ports.cpp:88-99,342-353 */
int cfd_advance(cfd_sim *s);            /* count one step done after a module sequence */

/* Host access (synchronising). */
int cfd_read_field(cfd_sim *s, int comp, double *host);        /* comp 0..5 Ex Ey Ez Hx Hy Hz */
int cfd_write_field(cfd_sim *s, int comp, const double *host);
/* cells [box0,box1) x [box2,box3) x [box4,box5) of one component, host
 * [i][j][k] k fastest: a sub-volume of Field::Ex...Hz without copying the
 * whole array (large grids) */
int cfd_read_region(cfd_sim *s, int comp, const int *box, double *host);
int cfd_read_id(cfd_sim *s, int16_t *host);                     /* Structure::ID */
int cfd_port_npts(cfd_sim *s, int port);
int cfd_port_points(cfd_sim *s, int port, int *ijk);
int cfd_port_samples(cfd_sim *s, int port, double *host);       /* [steps][npts][6] */
int cfd_port_gather(cfd_sim *s, int port, double *host);        /* [npts][6] current values:
                                                                   *Port::Update_Fields_on_Port Ports.cpp:88-99,342-353
                                                                   read back (port_point[ii].Ex ... .Hz) */
/* cfd_port_gather without any allocation: no counterpart in the reference.
 * For callers that cannot allocate device memory (tight serving loops,
 * harnesses that forbid hipMalloc).  Same values and layout as
 * cfd_port_gather; the values are staged through a fixed device array of
 * 1 Mi doubles per GPU (held until the process ends), so a port may have at
 * most 174762 points (error otherwise).  Calls on one GPU share that array:
 * the library serialises them within a process; calls from two processes on
 * one GPU would race.  Call with the simulation's device current. */
int cfd_port_gather_static(cfd_sim *s, int port, double *host);
int cfd_port_dft(cfd_sim *s, int port, double *host);           /* [npts][n_freq][6][2] */
int cfd_port_intensity(cfd_sim *s, int port, double *host);     /* [steps] sum of |E|^2 */
/* Spectral energy through a plane port, J/Hz, one value per DFT frequency:
 * 2 (stride dt)^2 Re sum_points (E_a H_b* - E_b H_a*) dA, (a, b, axis) right-
 * handed, dA the cell face normal to axis (0 x, 1 y, 2 z).  No reference
 * counterpart (Ports.cpp exports |E|^2 only); E and H at the point's own
 * indices (half a cell apart along the normal). */
int cfd_port_flux(cfd_sim *s, int port, int axis, double *host);
/* The same from DFT accumulators already on the host ([npts][nf][6][2], e.g.
 * a served request's CFD_RESULT_DFT): out[f] = scale Re sum_p (E_a H_b* -
 * E_b H_a*); scale = 2 (stride dt)^2 dA gives J/Hz.  No device needed. */
int cfd_dft_flux(const double *dft, int npts, int nf, int axis, double scale, double *out);
int cfd_gain_read(cfd_sim *s, double *host);                    /* [cells][4] N0..N3 */
/* Dispersive and gain state, cells in the reference's order (returns the
 * cell count; host may be NULL):
 * Drude-Lorentz, [cell][3 components][nD + 2 nL + 1]: the Drude currents,
 * the Lorentz currents, their previous values, the previous field
 * (DL_material JD*, JL*, JL*_1, E*_1, Drude_Lorentz.h:22-27, in 3-D);
 * gain, [cell][6]: polarisation x y z, then its previous values (Pa*,
 * Pa*_1 of FourLevelLaser.h:34,46; one transition in this 3-D model). */
int cfd_dl_state(cfd_sim *s, double *host);
int cfd_gain_state(cfd_sim *s, double *host);
/* i j k of each Drude-Lorentz (kind 0) or gain (kind 1) cell: DLM[].ix/jy
 * (Drude_Lorentz.h:25) and FourLevelLaser::location (FourLevelLaser.h:31). */
int cfd_special_cells(cfd_sim *s, int kind, int *ijk);
/* TF_SF::Einc and ::Hinc (TF_SF.h:57, filled by TF_SF.cpp:898-955); returns
 * the table length (einc, hinc may be NULL). */
int cfd_tfsf_wave(cfd_sim *s, double *einc, double *hinc);

/* Batches: n simulations of one shape (grid, CPML, periodic axes, TF/SF box
 * and waveform, ports, run length, precision) whose media, geometry and
 * dispersive cells may differ.  Each pass of a step is one launch over all
 * instances.  No reference counterpart (the reference runs one simulation
 * per process, main.cpp:115-343); each instance behaves as cfd_create of its
 * config followed by cfd_step, bit for bit. */
typedef struct cfd_batch cfd_batch;
int cfd_batch_create(const cfd_config *cfgs, int n, cfd_batch **out);
int cfd_batch_destroy(cfd_batch *b);
int cfd_batch_size(cfd_batch *b);
int cfd_batch_step(cfd_batch *b, int count);                 /* main.cpp:267-343 for every instance */
int cfd_batch_sync(cfd_batch *b);
/* a borrowed handle of instance i for the readers (cfd_read_field,
 * cfd_port_*, cfd_gain_read ...); it cannot be stepped or destroyed */
int cfd_batch_member(cfd_batch *b, int i, cfd_sim **out);
int cfd_batch_device_bytes(cfd_batch *b, size_t *bytes);
int cfd_batch_set_timing(cfd_batch *b, int on);
int cfd_batch_kernel_times(cfd_batch *b, char (*names)[32], double *ms, int *launches, int cap);

/* Packs: simulations of any grid size and physics (one precision) stepped
 * together, each pass of a step one launch over all members (the serving
 * loop's form).  No reference counterpart: the reference runs one simulation
 * per process (main.cpp:115-343).  A member keeps its handle and device
 * state (read it with the calls above once removed or the pack synchronised);
 * while in a pack it is stepped only by cfd_pack_step.  A simulation with a
 * running DFT joins at step 0 or at a multiple of 16 DFT strides.  Members
 * give the same results, bit for bit, as when stepped alone. */
typedef struct cfd_pack cfd_pack;
int cfd_pack_create(int precision, cfd_pack **out);
int cfd_pack_destroy(cfd_pack *p);                  /* members stay valid handles */
int cfd_pack_add(cfd_pack *p, cfd_sim *s);
int cfd_pack_remove(cfd_pack *p, cfd_sim *s);
int cfd_pack_size(cfd_pack *p);
/* non-finite field values of n members (counts[q] for sims[q]; the serving
 * loop's watchdog), counted after the pack's queued steps, one wait */
int cfd_pack_check_finite(cfd_pack *p, cfd_sim *const *sims, int n, int *counts);
int cfd_pack_step(cfd_pack *p, int count);          /* asynchronous; a finished member idles */
int cfd_pack_sync(cfd_pack *p);
int cfd_pack_query(cfd_pack *p);                    /* 1 done, 0 running */
int cfd_pack_set_timing(cfd_pack *p, int on);
int cfd_pack_kernel_times(cfd_pack *p, char (*names)[32], double *ms, int *launches, int cap);

/* Serving: requests (one cfd_config each) admitted first-come-first-served
 * while their device memory fits hbm_budget, then stepped in quanta of
 * `quantum` steps, each request on its own stream, or, below pack_cells
 * cells, as a member of a pack (one launch per pass over all of them; see
 * cfd_pack_*); a finished request's port results (|E|^2 series, DFT) are
 * copied to host memory and its device memory freed.  A scheduler thread
 * runs inside the server; the calls below are thread-safe.  No reference
 * counterpart (the reference runs one simulation per process,
 * main.cpp:115-343). */
#define CFD_REQ_QUEUED  0
#define CFD_REQ_RUNNING 1
#define CFD_REQ_DONE    2
#define CFD_REQ_FAILED  3
#define CFD_RESULT_INTENSITY 0   /* [steps] sum over the port of |E|^2 (Ports.cpp:512-522) */
#define CFD_RESULT_DFT       1   /* [points][n_freq][6][re,im] */
#define CFD_POLICY_FIFO     0   /* admit in arrival order */
#define CFD_POLICY_SHORTEST 1   /* admit the queued request with the least work (cells x steps) */
typedef struct {
    size_t hbm_budget;   /* bytes of device memory for running requests (0: 200 GiB) */
    int max_active;      /* requests running at once (0: 8) */
    int quantum;         /* steps per scheduling slice (0: 256) */
    int policy;          /* CFD_POLICY_* (0: FIFO) */
    double pack_cells;   /* requests with fewer cells are packed (0: none) */
    double step_cells;   /* stream requests stepping at once hold at most this many
                            cells together, the earliest admitted first; one always
                            steps (0: 2e7; < 0: no limit) */
} cfd_server_opts;
typedef struct cfd_server cfd_server;
int cfd_server_create(const cfd_server_opts *opts, cfd_server **out);
int cfd_server_destroy(cfd_server *s);                       /* stops; unfinished requests fail */
int cfd_server_submit(cfd_server *s, const cfd_config *cfg, long long *id);   /* copies cfg */
int cfd_server_poll(cfd_server *s, long long id, double *times);  /* returns CFD_REQ_*; times: submit, start, done (s) */
int cfd_server_wait(cfd_server *s, long long id, double timeout_s);  /* returns CFD_REQ_* */
int cfd_server_result(cfd_server *s, long long id, int port, int what, double *host, size_t *count);
const char *cfd_server_request_error(cfd_server *s, long long id);
int cfd_server_release(cfd_server *s, long long id);         /* drop a finished request's results */
int cfd_server_stats(cfd_server *s, long long *steps_launched, size_t *peak_bytes, int *peak_active, int *queued);
const char *cfd_server_last_error(void);

/* Host-side set-up only, no device: the parsed model the engine uploads
 * (Structure, CPML, TF_SF, port constructors).  Usable without a GPU. */
typedef struct cfd_model cfd_model;
int cfd_model_build(const cfd_config *cfg, cfd_model **out);
/* the device simulation of a built model; consumes the model (also on error) */
int cfd_create_from_model(cfd_model *model, int precision, cfd_sim **out);
int cfd_model_free(cfd_model *m);
const char *cfd_model_last_error(void);
/* A configuration's own fields, checked before any file is read (required
 * paths, counts against their arrays, port kinds, precision, periodic bits):
 * 0, or -1 with the message in cfd_model_last_error.  cfd_create,
 * cfd_model_build and cfd_server_submit apply the same check.  No reference
 * counterpart (the reference reads its inputs unchecked). */
int cfd_config_check(const cfd_config *cfg);
int cfd_model_dims(cfd_model *m, int *nxyz_nt, double *d);
int cfd_model_id(cfd_model *m, int16_t *host);                      /* Structure::ID */
/* named FP64 table; returns its length (call with out = NULL to size).
 * names: aE bE aH bH (Structure.cpp:352-357), stretch0..stretch11
 * (cExy cEyx cEzx cExz cEzy cEyz cHxy cHyx cHzx cHxz cHzy cHyz after CPML
 * kappa folding), pml_{E,H}{b,c}{x,y,z}{1,2} (CPML.cpp:139-229), Einc, Hinc
 * (TF_SF.cpp:898-955), dl (C1 C2 C3 C4[] C5[] C6[] kappa[] beta[] alpha[] xi[] gamma[]) */
int cfd_model_table(cfd_model *m, const char *name, double *out);
/* TF/SF correction lists: targets as comp i j k first count (6 ints), terms
 * as coef delay profile (3 doubles); returns the number of targets, *nterms
 * receives the number of terms.  h = 0 for E corrections, 1 for H. */
int cfd_model_tfsf(cfd_model *m, int h, int *targets, double *terms, int *nterms);
int cfd_model_port(cfd_model *m, int port, int *ijk);               /* returns point count */
/* A TF/SF file's incident tables alone, without a grid (TF_SF.cpp:898-955
 * with the file's lambda, dlambda, Nt and mode lines): Einc[n], Hinc[n] for
 * n = 0 .. max(Nt, nt) + 2 at time step dt; returns that length (einc, hinc
 * may be NULL), -1 with the message in cfd_model_last_error. */
int cfd_tfsf_tables(const char *tfsf, double dt, int nt, double *einc, double *hinc);
/* DL / gain cell lists in reference [i][j][k] flat order; returns count */
int cfd_model_cells(cfd_model *m, int gain, long long *flat);
/* Device bytes a simulation of this model will allocate at the given
 * precision (fields, IDs, tables, CPML psi, dispersive and gain state, TF/SF
 * lists, sources, port buffers, DFT accumulators, recorded samples), with a
 * small allowance; 0 for bad arguments.  No reference counterpart: the
 * serving loop's admission estimate. */
size_t cfd_model_device_bytes(cfd_model *m, int precision);

/* Image export of a plane port's maps, host only (OutImage.cpp; the
 * reference draws with CImg and saves with CImg::save_png).  Maps are
 * [nx][ny] with x along the image width (PlainPort::Export_EIntensity /
 * Export_ID layout, Ports.cpp:322-340); pixels are rgb[ny][nx][3]. */
/* OutImage::plot_save colour map (OutImage.cpp:140-202): updates rgb in place
 * (a NaN value leaves its pixel unchanged, as in the reference) */
int cfd_image_colormap(const double *map, int nx, int ny, double cmin, double cmax, unsigned char *rgb);
/* OutImage::structure_plot_save colour table (OutImage.cpp:44-53, 204-217);
 * an ID outside 0-9 is an error */
int cfd_image_materials(const short *ids, int nx, int ny, unsigned char *rgb);
/* CImg::save_png of those pixels (OutImage.cpp:199, 215): 8-bit RGB PNG */
int cfd_image_write_png(const char *path, int width, int height, const unsigned char *rgb);
const char *cfd_image_last_error(void);

/* Per-kernel timing of the last cfd_step call (HIP events on the sim's stream):
 * names[i] (<=32 chars) and total milliseconds; returns entries. */
int cfd_kernel_times(cfd_sim *s, char (*names)[32], double *ms, int *launches, int cap);
int cfd_set_timing(cfd_sim *s, int on);

#ifdef __cplusplus
}
#endif
#endif
