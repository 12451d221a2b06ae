// kernels.hip -- HIP kernels of the time step, for gfx950.
//
// All field kernels are bandwidth-bound stencils (about 0.7 flop per byte);
// nothing here is a dense contraction, so there is no MFMA.  Workgroups are
// TX x TY = 64 x 4 threads: one wavefront per x-row, so every field load and
// store of a wavefront is one contiguous, 128-byte-aligned run.  Tiles that
// do not touch a CPML slab skip the slab logic with a workgroup-uniform
// branch; tiles that do run the psi recursion fused into the half-step.
//
// Arithmetic keeps the reference's operand order, and the library is compiled
// with -ffp-contract=off, so the FP64 instantiation reproduces the reference's
// update (BasicFDTD.cpp, CPML.cpp, TF_SF.cpp) operation for operation.
#pragma once

#include <type_traits>

#include "engine.hpp"

namespace cfd {

enum { S_cExy, S_cEyx, S_cEzx, S_cExz, S_cEzy, S_cEyz, S_cHxy, S_cHyx, S_cHzx, S_cHxz, S_cHzy, S_cHyz };

__device__ __forceinline__ long long cell_index(int px, int py, int i, int j, int k)
{
    return (long long)i + (long long)px * ((long long)j + (long long)py * k);
}

// A half-step reads the updated component at its own cell once and writes it
// once; no other tile reads it in that pass.  When the fields do not fit in
// the Infinity Cache (NT), those loads and stores are non-temporal, so they
// stream past the caches and leave room for the neighbour values that
// adjacent tiles re-read (256^3: FP32 merged kernels 4-5 % faster, FP64 step
// 4.5 % faster).  When they fit, the cache keeps a pass's output for the next
// pass and plain accesses are faster (128^3: 3-10 %).
template <bool NT, class T>
__device__ __forceinline__ T ld_once(const T *p)
{
    if constexpr (NT) return __builtin_nontemporal_load(p);
    else return *p;
}
template <bool NT, class T>
__device__ __forceinline__ void st_once(T *p, T v)
{
    if constexpr (NT) __builtin_nontemporal_store(v, p);
    else *p = v;
}

// ---------------------------------------------------------------------------
// per-cell physics, shared by the grid kernels (module passes), the CPML
// shell and merged kernels (fused step) and the special pass.  Neighbour
// values arrive as arguments.
// ---------------------------------------------------------------------------

// A cell of the special pass (Drude-Lorentz or gain), in bitwise form: a
// short-circuit || becomes a branch, and the wait on the ID byte it needs
// then lands ahead of every later load (measured: the E CPML tiles issued
// their psi loads one memory round trip late).
template <class R>
__device__ __forceinline__ bool special_id(const Dev<R> &d, int id)
{
    return ((id == 2) & (d.ndl != 0)) | ((id == 5) & (d.ngain != 0));
}

// Coefficient of medium id.  With nsel (at most kSelMat media): a three-level
// select on the ID's bits over the values held in the kernel argument, so the
// update issues no load that depends on the ID byte (a table load is a second
// memory round trip after the field loads; 256^3 FP64 merged update kernels
// 273 -> 261 us with the selects).  Otherwise the table.
// SEL = false: always the table (kernels whose simulation view is read from
// device memory, where the selected values would be 32 more scalar loads).
template <bool SEL, class R>
__device__ __forceinline__ R mat_coef(const R (&v)[kSelMat], const R *tab, int nsel, int id)
{
    if (SEL && nsel) {
        const bool b0 = id & 1, b1 = id & 2, b2 = id & 4;
        const R v01 = b0 ? v[1] : v[0], v23 = b0 ? v[3] : v[2], v45 = b0 ? v[5] : v[4], v67 = b0 ? v[7] : v[6];
        const R v03 = b1 ? v23 : v01, v47 = b1 ? v67 : v45;
        return b2 ? v47 : v03;
    }
    return tab[id];
}
template <bool SEL = true, class R>
__device__ __forceinline__ R coef_aE(const Dev<R> &d, int id) { return mat_coef<SEL>(d.selE[0], d.aE, d.nsel, id); }
template <bool SEL = true, class R>
__device__ __forceinline__ R coef_bE(const Dev<R> &d, int id) { return mat_coef<SEL>(d.selE[1], d.bE, d.nsel, id); }
template <bool SEL = true, class R>
__device__ __forceinline__ R coef_aH(const Dev<R> &d, int id) { return mat_coef<SEL>(d.selH[0], d.aH, d.nsel, id); }
template <bool SEL = true, class R>
__device__ __forceinline__ R coef_bH(const Dev<R> &d, int id) { return mat_coef<SEL>(d.selH[1], d.bH, d.nsel, id); }

// CPML term of one axis at one cell.  The two slabs of an axis never overlap
// (the model rejects n_low + n_high >= N), so at most one holds the cell.
template <class R>
struct SlabTerm {
    R b, c, q1, q2;  // grading coefficients and the two psi values
    long long qi;    // psi index in the axis's arrays (the high slab's box follows the low one's)
    int side;        // -1: outside both slabs; 0: low side; 1: high side
};

// Locate the slab of each enabled axis (bit a of pm) holding cell (i, j, k)
// for field f (0: E, 1: H) and load its coefficients and both psi values, so
// the loads issue together with the caller's field loads.  E slabs start at
// index 1 on the low side, whose coefficient tables are indexed by the cell
// position (CPML.cpp:529-607); H slabs start at 0 (CPML.cpp:609-689).
// NT: psi read and written non-temporally (each psi pair is read and written
// once per pass; with streamed fields, keeping them out of the caches left
// room for the neighbour values: 256^3 FP32 update kernels 2.5-3 us faster)
template <class R, bool NT = false>
__device__ __forceinline__ void gather_slabs(const Dev<R> &d, int f, unsigned pm, int i, int j, int k,
                                             SlabTerm<R> (&T)[3])
{
    const int pos[3] = {i, j, k};
    const int N[3] = {d.Nx, d.Ny, d.Nz};
    const int o = f == 0 ? 1 : 0;
#pragma unroll
    for (int ax = 0; ax < 3; ax++) {
        T[ax].side = -1;
        if (!(pm & (1u << ax))) continue;
        const int n0 = d.npml[ax][0], n1 = d.npml[ax][1];
        const int x = pos[ax];
        int side = -1, l = 0, ci = 0;
        if (n0 && x - o >= 0 && x - o < n0) {
            side = 0; l = x - o; ci = x;
        } else if (n1 && x >= N[ax] - n1 && x < N[ax]) {
            side = 1; l = x - (N[ax] - n1); ci = l;
        }
        if (side < 0) continue;
        const int n = side ? n1 : n0;
        const long long qi = ax == 0 ? l + (long long)n * (j + (long long)d.py * k)
                           : ax == 1 ? i + (long long)d.px * (l + (long long)n * k)
                                     : i + (long long)d.px * (j + (long long)d.py * l);
        const int cb = side ? d.pofs[f][ax] : 0;
        const long long qb = side ? d.qofs[f][ax] : 0;
        const Pair<R> bc = d.pbc[f][ax][cb + ci];
        const Pair<R> *qp = &d.q[f][ax][qb + qi];
        const Pair<R> qq = {ld_once<NT>(&qp->a), ld_once<NT>(&qp->b)};
        T[ax].b = bc.a;
        T[ax].c = bc.b;
        T[ax].q1 = qq.a;
        T[ax].q2 = qq.b;
        T[ax].qi = qb + qi;  // index in the axis's psi array
        T[ax].side = side;
    }
}

template <class R, bool NT = false>
__device__ __forceinline__ void slab_store(const Dev<R> &d, int f, int ax, const SlabTerm<R> &t, R a1, R a2)
{
    Pair<R> *q = &d.q[f][ax][t.qi];
    st_once<NT>(&q->a, a1);
    st_once<NT>(&q->b, a2);
}

// Stretch factors of the interior box.  Outside the CPML slabs every entry of
// each stretch table the lean kernels read is the same value (the host checks
// this bit for bit over the box's index ranges); those kernels then take the
// twelve values as scalars and skip six table loads per cell.
template <class R>
struct LeanStretch {
    R s[12];
};
template <bool CS, class R>
__device__ __forceinline__ R lean_st(const Dev<R> &d, const LeanStretch<R> &c, int t, int idx)
{
    return CS ? c.s[t] : d.st[t][idx];
}

// Stretch value of table t at index idx along axis ax in a CPML-path cell.
// With CS (the tables are constant over the interior box), a tile whose slab
// mask pm leaves ax out lies in the interior box's range along ax, where the
// table holds that constant: no table load.
template <bool CS, class R>
__device__ __forceinline__ R st_at(const Dev<R> &d, const LeanStretch<R> *c, unsigned pm, int t, int ax, int idx)
{
    if constexpr (CS)
        if (!(pm & (1u << ax))) return c->s[t];
    return d.st[t][idx];
}

// H at one cell: BasicFDTD.cpp:67-81, then CPML.cpp:609-689 (x, then y, then
// z slab contributions, in the reference's order).  *_i1/_j1/_k1 are E at the
// +x/+y/+z neighbour.  Bit a of pm enables the slab test along axis a (a
// caller passes 0 for an axis none of its cells can be in).
// With CS, cs holds the interior stretch constants (st_at); NT as for
// gather_slabs.
template <class R, bool CS = false, bool NT = false, bool SEL = true>
__device__ __forceinline__ void h_cell(const Dev<R> &d, bool basic, unsigned pm, int i, int j, int k, long long p,
                                       R ex, R ey, R ez, R ex_j1, R ez_j1, R ex_k1, R ey_k1, R ey_i1, R ez_i1,
                                       R &hx, R &hy, R &hz, int id_loaded = -1, const LeanStretch<R> *cs = nullptr)
{
    // ID (unless the caller loaded it ahead of its field loads) and stretch
    // loads first, then the slab terms: every load of the cell is in flight
    // before the first wait
    const int id = basic ? (id_loaded >= 0 ? id_loaded : (int)d.id[p]) : 0;
    R sxz = 0, sxy = 0, syx = 0, syz = 0, szx = 0, szy = 0;
    if (basic) {
        sxz = st_at<CS>(d, cs, pm, S_cHxz, 1, j); sxy = st_at<CS>(d, cs, pm, S_cHxy, 2, k);
        syx = st_at<CS>(d, cs, pm, S_cHyx, 2, k); syz = st_at<CS>(d, cs, pm, S_cHyz, 0, i);
        szx = st_at<CS>(d, cs, pm, S_cHzx, 1, j); szy = st_at<CS>(d, cs, pm, S_cHzy, 0, i);
    }
    SlabTerm<R> T[3];
    gather_slabs<R, NT>(d, 1, pm, i, j, k, T);
    if (basic) {
        const R aH = coef_aH<SEL>(d, id), bH = coef_bH<SEL>(d, id);
        hx = aH * hx - bH * (ez_j1 - ez) * sxz + bH * (ey_k1 - ey) * sxy;
        hy = aH * hy - bH * (ex_k1 - ex) * syx + bH * (ez_i1 - ez) * syz;
        hz = aH * hz + bH * (ex_j1 - ex) * szx - bH * (ey_i1 - ey) * szy;
    }
    if (!pm) return;
    const R cH = d.cpmlH;
    const bool fixed = d.hsign_fixed;
    if (T[0].side >= 0) {  // x slab: Hy += psi(dEz/dx), Hz -= psi(dEy/dx)
        const R a1 = T[0].b * T[0].q1 + T[0].c * (ez_i1 - ez);
        const R a2 = T[0].b * T[0].q2 + T[0].c * (ey_i1 - ey);
        slab_store<R, NT>(d, 1, 0, T[0], a1, a2);
        hy = hy + a1 * cH;
        hz = hz - a2 * cH;
    }
    if (T[1].side >= 0) {  // y slab: Hz -/+ psi(dEx/dy), Hx +/- psi(dEz/dy)
        const R a1 = T[1].b * T[1].q1 + T[1].c * (ex_j1 - ex);
        const R a2 = T[1].b * T[1].q2 + T[1].c * (ez_j1 - ez);
        slab_store<R, NT>(d, 1, 1, T[1], a1, a2);
        hz = fixed ? hz + a1 * cH : hz - a1 * cH;
        hx = fixed ? hx - a2 * cH : hx + a2 * cH;
    }
    if (T[2].side >= 0) {  // z slab: Hx -/+ psi(dEy/dz), Hy +/- psi(dEx/dz)
        const R a1 = T[2].b * T[2].q1 + T[2].c * (ey_k1 - ey);
        const R a2 = T[2].b * T[2].q2 + T[2].c * (ex_k1 - ex);
        slab_store<R, NT>(d, 1, 2, T[2], a1, a2);
        hx = fixed ? hx + a1 * cH : hx - a1 * cH;
        hy = fixed ? hy - a2 * cH : hy + a2 * cH;
    }
}

// E at one cell: four-level gain polarisation from E^n, BasicFDTD.cpp:25-45
// (ID 4 zeroed, ID 2 left to the ADE), the Drude-Lorentz ADE for ID 2
// (Drude_Lorentz.cpp:172-206, all three components, C1/C5 per F9/F10), the
// gain current and FP64 populations for ID 5, then CPML.cpp:529-607.
// *_im/_jm/_km are H (already updated this step) at the -x/-y/-z neighbour;
// pm as for h_cell.
template <class R, bool CS = false, bool NT = false, bool SEL = true>
__device__ __forceinline__ void e_cell(const Dev<R> &d, unsigned parts, unsigned pm, int i, int j, int k, long long p,
                                       int id, R hx, R hy, R hz, R hx_jm, R hz_jm, R hx_km, R hy_km, R hy_im,
                                       R hz_im, R (&E)[3], bool skip_special = false, const LeanStretch<R> *cs = nullptr)
{
    // stretch values and slab terms loaded ahead of every store below
    R sxz = 0, sxy = 0, syx = 0, syz = 0, szy = 0, szx = 0;
    if (parts & (E_BASIC | E_DL)) {
        sxz = st_at<CS>(d, cs, pm, S_cExz, 1, j); sxy = st_at<CS>(d, cs, pm, S_cExy, 2, k);
        syx = st_at<CS>(d, cs, pm, S_cEyx, 2, k); syz = st_at<CS>(d, cs, pm, S_cEyz, 0, i);
        szy = st_at<CS>(d, cs, pm, S_cEzy, 0, i); szx = st_at<CS>(d, cs, pm, S_cEzx, 1, j);
    }
    SlabTerm<R> T[3];
    gather_slabs<R, NT>(d, 0, pm, i, j, k, T);
    const unsigned gparts = parts & (E_GAINPOL | E_GAINE | E_GAINPOP);
    const bool gain = gparts && d.ngain && id == 5;
    const bool dl = (parts & E_DL) && d.ndl && id == 2;
    const R En[3] = {E[0], E[1], E[2]};  // E^n, kept for the gain populations
    int gs = -1;
    if (gain) {
        gs = d.aux[p];
        if (parts & E_GAINPOL) {
            // CPa3 (N1 - N2) formed in FP64 and rounded once: CPa3 alone is
            // below FP32's smallest normal number (about 7e-39 at 20 nm cells)
            const double dN = d.gN[1LL * d.ngain + gs] - d.gN[2LL * d.ngain + gs];
            const R cdN = (R)(d.g_CPa3 * dN);
            for (int c = 0; c < 3; c++) {
                const long long o = (long long)c * d.ngain + gs;
                const R Pn = d.gP[o];
                d.gP[o] = d.CPa1 * Pn + d.CPa2 * d.gP1[o] + cdN * E[c];
                d.gP1[o] = Pn;
                if (!(parts & E_GAINPOP)) d.gEprev[o] = E[c];
            }
        }
    }
    if (parts & E_BASIC) {  // branch-free: ID 4 zeroed, ID 2 left as is
        const R aE = coef_aE<SEL>(d, id), bE = coef_bE<SEL>(d, id);
        const R u[3] = {aE * E[0] + bE * (hz - hz_jm) * sxz - bE * (hy - hy_km) * sxy,
                        aE * E[1] + bE * (hx - hx_km) * syx - bE * (hz - hz_im) * syz,
                        aE * E[2] + bE * (hy - hy_im) * szy - bE * (hx - hx_jm) * szx};
        for (int c = 0; c < 3; c++) E[c] = id == 4 ? (R)0 : id == 2 ? E[c] : u[c];
    }
    if (dl) {
        const int s = d.aux[p];
        const R dA[3] = {hz - hz_jm, hx - hx_km, hy - hy_im};
        const R cA[3] = {sxz, syx, szy};
        const R dB[3] = {hy - hy_km, hz - hz_im, hx - hx_jm};
        const R cB[3] = {sxy, syz, szx};
        const int nD = d.nD, nL = d.nL;
        const long long ndl = d.ndl, vs = 3 * ndl;
        for (int c = 0; c < 3; c++) {
            R *JD = d.dlst + (long long)c * ndl + s;  // variable v at JD + v*3*ndl
            R *JL = JD + (long long)nD * vs;
            R *JL1 = JL + (long long)nL * vs;
            R *E1 = JL1 + (long long)nL * vs;
            const R temp = E[c];
            const R e1 = *E1;
            R t1 = 0, t2 = 0;
            for (int q = 0; q < nD; q++) t1 = t1 + d.C4[q] * JD[q * vs];
            for (int q = 0; q < nL; q++) t2 = t2 + d.C5[q] * JL[q * vs] + d.C6[q] * JL1[q * vs];
            const R en = d.C1 * E[c] + d.C2 * e1 + d.C3 * dA[c] * cA[c] - d.C3 * dB[c] * cB[c] - t1 - t2;
            E[c] = en;
            for (int q = 0; q < nD; q++) JD[q * vs] = d.kap[q] * JD[q * vs] + d.beta[q] * (en + temp);
            for (int q = 0; q < nL; q++) {
                const R t0 = JL[q * vs];
                JL[q * vs] = d.alpha[q] * JL[q * vs] + d.xi[q] * JL1[q * vs] + d.gam[q] * (en - e1) / (R)2.0 / d.dt;
                JL1[q * vs] = t0;
            }
            *E1 = temp;
        }
    }
    if (gain && (gparts & (E_GAINE | E_GAINPOP))) {
        const R coef = coef_bE<SEL>(d, id) / d.dt;
        double dot = 0.0;
        for (int c = 0; c < 3; c++) {
            const long long o = (long long)c * d.ngain + gs;
            const R dP = d.gP[o] - d.gP1[o];
            if (parts & E_GAINE) E[c] = E[c] - coef * dP;
            const R eprev = (parts & E_GAINPOL) ? En[c] : d.gEprev[o];
            dot = dot + (double)(E[c] + eprev) * (double)dP;
        }
        if (parts & E_GAINPOP) {  // rate equations, Crank-Nicolson decay terms, FP64
            const long long ng = d.ngain;
            const double Rs = dot / (2.0 * d.g_dt) / d.g_hw;
            const double N0 = d.gN[gs], N1 = d.gN[ng + gs], N2 = d.gN[2 * ng + gs], N3 = d.gN[3 * ng + gs];
            const double h32 = d.g_h32, h21 = d.g_h21, h10 = d.g_h10, dt = d.g_dt;
            const double n3 = (N3 * (1.0 - h32) + d.g_dtWp * N0) / (1.0 + h32);
            const double n2 = (N2 * (1.0 - h21) + h32 * (N3 + n3) + dt * Rs) / (1.0 + h21);
            const double n1 = (N1 * (1.0 - h10) + h21 * (N2 + n2) - dt * Rs) / (1.0 + h10);
            d.gN[3 * ng + gs] = n3;
            d.gN[2 * ng + gs] = n2;
            d.gN[ng + gs] = n1;
            d.gN[gs] = d.g_Ntot - n1 - n2 - n3;
        }
    }
    if (!pm) return;
    // with skip_special, the special pass's cells keep their psi (that pass
    // updates it); tested here, after every load above has been issued
    const bool psi_store = !(skip_special && special_id(d, id));
    const R cE = d.cpmlE;
    if (T[0].side >= 0) {  // x slab: Ey -= psi(dHz/dx), Ez += psi(dHy/dx)
        const R a1 = T[0].b * T[0].q1 + T[0].c * (hz - hz_im);
        const R a2 = T[0].b * T[0].q2 + T[0].c * (hy - hy_im);
        if (psi_store) slab_store<R, NT>(d, 0, 0, T[0], a1, a2);
        E[1] = E[1] - a1 * cE;
        E[2] = E[2] + a2 * cE;
    }
    if (T[1].side >= 0) {  // y slab: Ez -= psi(dHx/dy), Ex += psi(dHz/dy)
        const R a1 = T[1].b * T[1].q1 + T[1].c * (hx - hx_jm);
        const R a2 = T[1].b * T[1].q2 + T[1].c * (hz - hz_jm);
        if (psi_store) slab_store<R, NT>(d, 0, 1, T[1], a1, a2);
        E[2] = E[2] - a1 * cE;
        E[0] = E[0] + a2 * cE;
    }
    if (T[2].side >= 0) {  // z slab: Ex -= psi(dHy/dz), Ey += psi(dHx/dz)
        const R a1 = T[2].b * T[2].q1 + T[2].c * (hy - hy_km);
        const R a2 = T[2].b * T[2].q2 + T[2].c * (hx - hx_km);
        if (psi_store) slab_store<R, NT>(d, 0, 2, T[2], a1, a2);
        E[0] = E[0] - a1 * cE;
        E[1] = E[1] + a2 * cE;
    }
}

// tile-level CPML membership (workgroup-uniform)
template <class R>
__device__ __forceinline__ bool tile_in_pml_xy(const Dev<R> &d, int x0, int y0, bool efield)
{
    const int o = efield ? 1 : 0;  // E slabs start at index 1 on the low side
    return (d.npml[0][0] && x0 < d.npml[0][0] + o) || (d.npml[0][1] && x0 + TX > d.Nx - d.npml[0][1]) ||
           (d.npml[1][0] && y0 < d.npml[1][0] + o) || (d.npml[1][1] && y0 + TY > d.Ny - d.npml[1][1]);
}
template <class R>
__device__ __forceinline__ bool plane_in_pml_z(const Dev<R> &d, int k, bool efield)
{
    const int o = efield ? 1 : 0;
    return (d.npml[2][0] && k < d.npml[2][0] + o) || (d.npml[2][1] && k >= d.Nz - d.npml[2][1]);
}

// ---------------------------------------------------------------------------
// grid kernels: one thread per cell, workgroup = 64 x 4 tile of one plane.
// Used for the single-module passes (reference call granularity).
// ---------------------------------------------------------------------------
// grid kernels: one thread per cell, workgroups in linear tile order (x
// fastest, then y, then z): all XCDs sweep the same planes together.
struct TileGrid {
    int gx, gy, gz;
};
__device__ __forceinline__ bool tile_of(const TileGrid &g, int &bx, int &by, int &bz)
{
    const int b = blockIdx.x;
    if (b >= g.gx * g.gy * g.gz) return false;
    bx = b % g.gx;
    by = (b / g.gx) % g.gy;
    bz = b / (g.gx * g.gy);
    return true;
}

template <class R>
__global__ __launch_bounds__(TX *TY) void k_update_h(Dev<R> d, unsigned parts, TileGrid g)
{
    int bx, by, k;
    if (!tile_of(g, bx, by, k)) return;
    const int x0 = bx * TX, y0 = by * TY;
    const int i = x0 + threadIdx.x, j = y0 + threadIdx.y;
    const bool pml = (parts & H_PML) && d.has_pml && (tile_in_pml_xy(d, x0, y0, false) || plane_in_pml_z(d, k, false));
    if (!(parts & H_BASIC) && !pml) return;
    if (i >= d.Nx || j >= d.Ny) return;
    const long long p = cell_index(d.px, d.py, i, j, k);
    const long long pi1 = (d.pbx && i + 1 == d.Nx) ? p + 1 - d.Nx : p + 1;
    const long long pj1 = (d.pby && j + 1 == d.Ny) ? p + d.px - (long long)d.Ny * d.px : p + d.px;
    const long long pk1 = p + d.slab;
    R hx = d.F[3][p], hy = d.F[4][p], hz = d.F[5][p];
    h_cell(d, (parts & H_BASIC) != 0, pml ? 7u : 0u, i, j, k, p, d.F[0][p], d.F[1][p], d.F[2][p], d.F[0][pj1], d.F[2][pj1],
           d.F[0][pk1], d.F[1][pk1], d.F[1][pi1], d.F[2][pi1], hx, hy, hz);
    st_once<false>(&d.F[3][p], hx);
    st_once<false>(&d.F[4][p], hy);
    st_once<false>(&d.F[5][p], hz);
}

template <class R>
__global__ __launch_bounds__(TX *TY) void k_update_e(Dev<R> d, unsigned parts, TileGrid g)
{
    int bx, by, k;
    if (!tile_of(g, bx, by, k)) return;
    const int x0 = bx * TX, y0 = by * TY;
    const int i = x0 + threadIdx.x, j = y0 + threadIdx.y;
    const bool pml = (parts & E_PML) && d.has_pml && (tile_in_pml_xy(d, x0, y0, true) || plane_in_pml_z(d, k, true));
    if (!(parts & ~E_PML) && !pml) return;
    if (i < d.ie0 || i >= d.Nx || j < d.je0 || j >= d.Ny || k < 1 || k >= d.Nz) return;
    const long long p = cell_index(d.px, d.py, i, j, k);
    const int id = d.id[p];
    const bool gain = (parts & (E_GAINPOL | E_GAINE | E_GAINPOP)) && d.ngain && id == 5;
    const bool dl = (parts & E_DL) && d.ndl && id == 2;
    if (!(parts & (E_BASIC | E_PML)) && !gain && !dl) return;
    const long long pim = (i == 0) ? p + d.Nx - 1 : p - 1;
    const long long pjm = (j == 0) ? p + (long long)(d.Ny - 1) * d.px : p - d.px;
    const long long pkm = p - d.slab;
    R E[3] = {d.F[0][p], d.F[1][p], d.F[2][p]};
    e_cell(d, parts, pml ? 7u : 0u, i, j, k, p, id, d.F[3][p], d.F[4][p], d.F[5][p], d.F[3][pjm], d.F[5][pjm], d.F[3][pkm],
           d.F[4][pkm], d.F[4][pim], d.F[5][pim], E);
    st_once<false>(&d.F[0][p], E[0]);
    st_once<false>(&d.F[1][p], E[1]);
    st_once<false>(&d.F[2][p], E[2]);
}

// ---------------------------------------------------------------------------
// Fused-step kernels.  A half-step is split by cell class so that the hot
// path stays lean (few registers, high occupancy):
//   lean     the interior box (no CPML slab): plain Yee stencil; on the E side
//            Drude-Lorentz / gain cells are skipped (left to the special pass)
//   shell    the CPML slab boxes: full physics (stencil + psi + ADE + gain)
//   special  interior Drude-Lorentz and gain cells, from a compact list
// Every cell belongs to exactly one class, and within a half-step a cell's
// update reads only the other field, so the passes are independent and each
// cell's own sequence of operations is the reference's.
// ---------------------------------------------------------------------------
struct Box {
    int i0, i1, j0, j1, k0, k1;  // cell ranges [lo, hi)
    int tw, th;                  // tile width (x) and height (y), tw * th = 256
    int xbase, gx, gy, bstart;   // x tiles start at xbase (aligned); block offset
    int pml;                     // bit a: cells of the box may lie in a CPML slab along axis a
};

// the shell boxes of one half-step, passed by value (kernel-argument memory)
constexpr int kMaxBoxes = 6;
struct BoxSet {
    Box b[kMaxBoxes];
    int n;
};

// locate this workgroup's box, tile and cell; returns false for lanes outside
__device__ __forceinline__ bool box_cell(const BoxSet &S, int &i, int &j, int &k, int &bi)
{
    int q = 0;
#pragma unroll
    for (int s = 1; s < kMaxBoxes; s++)
        if (s < S.n && (int)blockIdx.x >= S.b[s].bstart) q = s;
    const Box &B = S.b[q];
    const int t = blockIdx.x - B.bstart;
    const int bx = t % B.gx, by = (t / B.gx) % B.gy;
    k = B.k0 + t / (B.gx * B.gy);
    const int lin = threadIdx.y * TX + threadIdx.x;
    i = B.xbase + bx * B.tw + lin % B.tw;
    j = B.j0 + by * B.th + lin / B.tw;
    bi = q;
    return i >= B.i0 && i < B.i1 && j < B.j1;
}


// interior H: BasicFDTD.cpp:67-81 only
template <class R, bool CS, bool NT>
__global__ __launch_bounds__(TX *TY) void k_lean_h(Dev<R> d, Box B, LeanStretch<R> c)
{
    const int t = blockIdx.x;
    const int bx = t % B.gx, by = (t / B.gx) % B.gy, k = B.k0 + t / (B.gx * B.gy);
    const int i = B.xbase + bx * TX + threadIdx.x, j = B.j0 + by * TY + threadIdx.y;
    if (i < B.i0 || i >= B.i1 || j >= B.j1) return;
    const long long p = cell_index(d.px, d.py, i, j, k);
    const long long pi1 = (d.pbx && i + 1 == d.Nx) ? p + 1 - d.Nx : p + 1;
    const long long pj1 = (d.pby && j + 1 == d.Ny) ? p + d.px - (long long)d.Ny * d.px : p + d.px;
    const long long pk1 = p + d.slab;
    const int id = d.id[p];  // first: the material lookup then overlaps the field loads
    const R ex = d.F[0][p], ey = d.F[1][p], ez = d.F[2][p];
    R hx = ld_once<NT>(&d.F[3][p]), hy = ld_once<NT>(&d.F[4][p]), hz = ld_once<NT>(&d.F[5][p]);
    const R aH = coef_aH(d, id), bH = coef_bH(d, id);
    hx = aH * hx - bH * (d.F[2][pj1] - ez) * lean_st<CS>(d, c, S_cHxz, j) + bH * (d.F[1][pk1] - ey) * lean_st<CS>(d, c, S_cHxy, k);
    hy = aH * hy - bH * (d.F[0][pk1] - ex) * lean_st<CS>(d, c, S_cHyx, k) + bH * (d.F[2][pi1] - ez) * lean_st<CS>(d, c, S_cHyz, i);
    hz = aH * hz + bH * (d.F[0][pj1] - ex) * lean_st<CS>(d, c, S_cHzx, j) - bH * (d.F[1][pi1] - ey) * lean_st<CS>(d, c, S_cHzy, i);
    st_once<NT>(&d.F[3][p], hx);
    st_once<NT>(&d.F[4][p], hy);
    st_once<NT>(&d.F[5][p], hz);
}

// interior E: BasicFDTD.cpp:25-45 for plain cells (ID 4 zeroed; ID 2 and, with a
// gain medium, ID 5 are the special pass's)
template <class R, bool CS, bool NT>
__global__ __launch_bounds__(TX *TY) void k_lean_e(Dev<R> d, Box B, LeanStretch<R> c, int skip5)
{
    const int t = blockIdx.x;
    const int bx = t % B.gx, by = (t / B.gx) % B.gy, k = B.k0 + t / (B.gx * B.gy);
    const int i = B.xbase + bx * TX + threadIdx.x, j = B.j0 + by * TY + threadIdx.y;
    if (i < B.i0 || i >= B.i1 || j >= B.j1) return;
    const long long p = cell_index(d.px, d.py, i, j, k);
    const long long pim = (i == 0) ? p + d.Nx - 1 : p - 1;
    const long long pjm = (j == 0) ? p + (long long)(d.Ny - 1) * d.px : p - d.px;
    const long long pkm = p - d.slab;
    // Branch-free on the material: every value is used on every path, so the
    // field loads issue together with the ID load instead of after it (a
    // branch on the ID lets the compiler sink them past a wait).  Cells of the
    // special pass store their own value back unchanged.
    const int id = d.id[p];
    const R hx = d.F[3][p], hy = d.F[4][p], hz = d.F[5][p];
    const R ex = ld_once<NT>(&d.F[0][p]), ey = ld_once<NT>(&d.F[1][p]), ez = ld_once<NT>(&d.F[2][p]);
    const R hz_jm = d.F[5][pjm], hy_km = d.F[4][pkm], hx_km = d.F[3][pkm];
    const R hz_im = d.F[5][pim], hy_im = d.F[4][pim], hx_jm = d.F[3][pjm];
    const R aE = coef_aE(d, id), bE = coef_bE(d, id);
    R e0 = aE * ex + bE * (hz - hz_jm) * lean_st<CS>(d, c, S_cExz, j) - bE * (hy - hy_km) * lean_st<CS>(d, c, S_cExy, k);
    R e1 = aE * ey + bE * (hx - hx_km) * lean_st<CS>(d, c, S_cEyx, k) - bE * (hz - hz_im) * lean_st<CS>(d, c, S_cEyz, i);
    R e2 = aE * ez + bE * (hy - hy_im) * lean_st<CS>(d, c, S_cEzy, i) - bE * (hx - hx_jm) * lean_st<CS>(d, c, S_cEzx, j);
    const bool keep = (id == 2) | ((skip5 != 0) & (id == 5));
    e0 = keep ? ex : id == 4 ? (R)0 : e0;
    e1 = keep ? ey : id == 4 ? (R)0 : e1;
    e2 = keep ? ez : id == 4 ? (R)0 : e2;
    st_once<NT>(&d.F[0][p], e0);
    st_once<NT>(&d.F[1][p], e1);
    st_once<NT>(&d.F[2][p], e2);
}

// CPML slab boxes: full physics, slab tests limited to the box's axes
template <class R, bool NT>
__global__ __launch_bounds__(TX *TY) void k_shell_h(Dev<R> d, BoxSet S)
{
    int i, j, k, bi;
    if (!box_cell(S, i, j, k, bi)) return;
    const long long p = cell_index(d.px, d.py, i, j, k);
    const long long pi1 = (d.pbx && i + 1 == d.Nx) ? p + 1 - d.Nx : p + 1;
    const long long pj1 = (d.pby && j + 1 == d.Ny) ? p + d.px - (long long)d.Ny * d.px : p + d.px;
    const long long pk1 = p + d.slab;
    const int id = d.id[p];  // first: the material lookup then overlaps the field loads
    R hx = d.F[3][p], hy = d.F[4][p], hz = d.F[5][p];
    h_cell<R, false, NT>(d, true, (unsigned)S.b[bi].pml, i, j, k, p, d.F[0][p], d.F[1][p], d.F[2][p], d.F[0][pj1], d.F[2][pj1],
           d.F[0][pk1], d.F[1][pk1], d.F[1][pi1], d.F[2][pi1], hx, hy, hz, id);
    st_once<NT>(&d.F[3][p], hx);
    st_once<NT>(&d.F[4][p], hy);
    st_once<NT>(&d.F[5][p], hz);
}

template <class R>
__device__ __forceinline__ void e_full(const Dev<R> &d, int i, int j, int k, unsigned pm)
{
    const long long p = cell_index(d.px, d.py, i, j, k);
    const long long pim = (i == 0) ? p + d.Nx - 1 : p - 1;
    const long long pjm = (j == 0) ? p + (long long)(d.Ny - 1) * d.px : p - d.px;
    const long long pkm = p - d.slab;
    R E[3] = {d.F[0][p], d.F[1][p], d.F[2][p]};
    e_cell(d, E_ALL, pm, i, j, k, p, d.id[p], d.F[3][p], d.F[4][p], d.F[5][p], d.F[3][pjm], d.F[5][pjm],
           d.F[3][pkm], d.F[4][pkm], d.F[4][pim], d.F[5][pim], E);
    st_once<false>(&d.F[0][p], E[0]);
    st_once<false>(&d.F[1][p], E[1]);
    st_once<false>(&d.F[2][p], E[2]);
}

// e_full with the pole counts fixed at compile time (ND Drude, NL Lorentz
// poles): the same operations in the same order as e_cell(E_ALL), but every
// state load of the cell (ADE currents and E^{n-1}, or the gain polarisation
// and populations) is issued before the first state store.  In the generic
// form each component's state is stored before the next component's is
// loaded, an order the compiler must keep (same arrays), so those loads went
// out one memory round trip at a time.
template <class R, int ND, int NL>
__device__ __forceinline__ void e_special(const Dev<R> &d, int i, int j, int k, unsigned pm)
{
    const long long p = cell_index(d.px, d.py, i, j, k);
    const long long pim = (i == 0) ? p + d.Nx - 1 : p - 1;
    const long long pjm = (j == 0) ? p + (long long)(d.Ny - 1) * d.px : p - d.px;
    const long long pkm = p - d.slab;
    const int id = d.id[p];
    const int s = d.aux[p];  // state slot of this Drude-Lorentz or gain cell
    R E[3] = {d.F[0][p], d.F[1][p], d.F[2][p]};
    const R hx = d.F[3][p], hy = d.F[4][p], hz = d.F[5][p];
    const R hx_jm = d.F[3][pjm], hz_jm = d.F[5][pjm], hx_km = d.F[3][pkm];
    const R hy_km = d.F[4][pkm], hy_im = d.F[4][pim], hz_im = d.F[5][pim];
    const R sxz = d.st[S_cExz][j], sxy = d.st[S_cExy][k], syx = d.st[S_cEyx][k];
    const R syz = d.st[S_cEyz][i], szy = d.st[S_cEzy][i], szx = d.st[S_cEzx][j];
    SlabTerm<R> T[3];
    gather_slabs(d, 0, pm, i, j, k, T);
    const bool gain = d.ngain && id == 5;
    const bool dl = (ND + NL > 0) && d.ndl && id == 2;

    // every state load first
    constexpr int nd = ND > 0 ? ND : 1, nl = NL > 0 ? NL : 1;
    R jd[3][nd], jl[3][nl], jl1[3][nl], e1[3];
    const long long ndl = d.ndl, vs = 3 * ndl;
    if (dl) {
        for (int c = 0; c < 3; c++) {
            const R *JD = d.dlst + (long long)c * ndl + s;  // variable v at JD + v*3*ndl
            const R *JL = JD + (long long)ND * vs;
            const R *JL1 = JL + (long long)NL * vs;
            for (int q = 0; q < ND; q++) jd[c][q] = JD[q * vs];
            for (int q = 0; q < NL; q++) jl[c][q] = JL[q * vs];
            for (int q = 0; q < NL; q++) jl1[c][q] = JL1[q * vs];
            e1[c] = JL1[(long long)NL * vs];
        }
    }
    const long long ng = d.ngain;
    R P[3] = {0, 0, 0}, P1[3] = {0, 0, 0};
    double N[4] = {0, 0, 0, 0};
    if (gain) {
        for (int c = 0; c < 3; c++) {
            P[c] = d.gP[(long long)c * ng + s];
            P1[c] = d.gP1[(long long)c * ng + s];
        }
        for (int q = 0; q < 4; q++) N[q] = d.gN[(long long)q * ng + s];
    }

    const R En[3] = {E[0], E[1], E[2]};  // E^n, kept for the gain populations
    R Pnew[3] = {0, 0, 0};
    if (gain) {  // polarisation P^{n+1} from E^n
        const R cdN = (R)(d.g_CPa3 * (N[1] - N[2]));  // FP64 product, as in e_cell
        for (int c = 0; c < 3; c++) {
            const long long o = (long long)c * ng + s;
            Pnew[c] = d.CPa1 * P[c] + d.CPa2 * P1[c] + cdN * E[c];
            d.gP[o] = Pnew[c];
            d.gP1[o] = P[c];
        }
    }
    {  // basic update: ID 2 left to the ADE below
        const R aE = coef_aE(d, id), bE = coef_bE(d, id);
        const R u[3] = {aE * E[0] + bE * (hz - hz_jm) * sxz - bE * (hy - hy_km) * sxy,
                        aE * E[1] + bE * (hx - hx_km) * syx - bE * (hz - hz_im) * syz,
                        aE * E[2] + bE * (hy - hy_im) * szy - bE * (hx - hx_jm) * szx};
        for (int c = 0; c < 3; c++) E[c] = id == 4 ? (R)0 : id == 2 ? E[c] : u[c];
    }
    if (dl) {
        const R dA[3] = {hz - hz_jm, hx - hx_km, hy - hy_im};
        const R cA[3] = {sxz, syx, szy};
        const R dB[3] = {hy - hy_km, hz - hz_im, hx - hx_jm};
        const R cB[3] = {sxy, syz, szx};
        for (int c = 0; c < 3; c++) {
            R *JD = d.dlst + (long long)c * ndl + s;
            R *JL = JD + (long long)ND * vs;
            R *JL1 = JL + (long long)NL * vs;
            R *E1 = JL1 + (long long)NL * vs;
            const R temp = E[c];
            R t1 = 0, t2 = 0;
            for (int q = 0; q < ND; q++) t1 = t1 + d.C4[q] * jd[c][q];
            for (int q = 0; q < NL; q++) t2 = t2 + d.C5[q] * jl[c][q] + d.C6[q] * jl1[c][q];
            const R en = d.C1 * E[c] + d.C2 * e1[c] + d.C3 * dA[c] * cA[c] - d.C3 * dB[c] * cB[c] - t1 - t2;
            E[c] = en;
            for (int q = 0; q < ND; q++) JD[q * vs] = d.kap[q] * jd[c][q] + d.beta[q] * (en + temp);
            for (int q = 0; q < NL; q++) {
                JL[q * vs] = d.alpha[q] * jl[c][q] + d.xi[q] * jl1[c][q] + d.gam[q] * (en - e1[c]) / (R)2.0 / d.dt;
                JL1[q * vs] = jl[c][q];
            }
            *E1 = temp;
        }
    }
    if (gain) {  // E -= (bE/dt) dP, then the populations (FP64)
        const R coef = coef_bE(d, id) / d.dt;
        double dot = 0.0;
        for (int c = 0; c < 3; c++) {
            const R dP = Pnew[c] - P[c];
            E[c] = E[c] - coef * dP;
            dot = dot + (double)(E[c] + En[c]) * (double)dP;
        }
        const double Rs = dot / (2.0 * d.g_dt) / d.g_hw;
        const double h32 = d.g_h32, h21 = d.g_h21, h10 = d.g_h10, dt = d.g_dt;
        const double n3 = (N[3] * (1.0 - h32) + d.g_dtWp * N[0]) / (1.0 + h32);
        const double n2 = (N[2] * (1.0 - h21) + h32 * (N[3] + n3) + dt * Rs) / (1.0 + h21);
        const double n1 = (N[1] * (1.0 - h10) + h21 * (N[2] + n2) - dt * Rs) / (1.0 + h10);
        d.gN[3 * ng + s] = n3;
        d.gN[2 * ng + s] = n2;
        d.gN[ng + s] = n1;
        d.gN[s] = d.g_Ntot - n1 - n2 - n3;
    }
    if (pm) {  // CPML slab terms, the special pass's cells keep them here
        const R cE = d.cpmlE;
        if (T[0].side >= 0) {
            const R a1 = T[0].b * T[0].q1 + T[0].c * (hz - hz_im);
            const R a2 = T[0].b * T[0].q2 + T[0].c * (hy - hy_im);
            slab_store(d, 0, 0, T[0], a1, a2);
            E[1] = E[1] - a1 * cE;
            E[2] = E[2] + a2 * cE;
        }
        if (T[1].side >= 0) {
            const R a1 = T[1].b * T[1].q1 + T[1].c * (hx - hx_jm);
            const R a2 = T[1].b * T[1].q2 + T[1].c * (hz - hz_jm);
            slab_store(d, 0, 1, T[1], a1, a2);
            E[2] = E[2] - a1 * cE;
            E[0] = E[0] + a2 * cE;
        }
        if (T[2].side >= 0) {
            const R a1 = T[2].b * T[2].q1 + T[2].c * (hy - hy_km);
            const R a2 = T[2].b * T[2].q2 + T[2].c * (hx - hx_km);
            slab_store(d, 0, 2, T[2], a1, a2);
            E[0] = E[0] - a1 * cE;
            E[1] = E[1] + a2 * cE;
        }
    }
    d.F[0][p] = E[0];
    d.F[1][p] = E[1];
    d.F[2][p] = E[2];
}

// CPML shell, E: stencil and psi only.  Drude-Lorentz and gain cells of the
// shell belong to the special pass, which also runs their psi updates.
template <class R, bool NT>
__global__ __launch_bounds__(TX *TY) void k_shell_e(Dev<R> d, BoxSet S)
{
    int i, j, k, bi;
    if (!box_cell(S, i, j, k, bi)) return;
    const long long p = cell_index(d.px, d.py, i, j, k);
    const long long pim = (i == 0) ? p + d.Nx - 1 : p - 1;
    const long long pjm = (j == 0) ? p + (long long)(d.Ny - 1) * d.px : p - d.px;
    const long long pkm = p - d.slab;
    const int id = d.id[p];
    R E[3] = {ld_once<NT>(&d.F[0][p]), ld_once<NT>(&d.F[1][p]), ld_once<NT>(&d.F[2][p])};
    const R hx = d.F[3][p], hy = d.F[4][p], hz = d.F[5][p];
    const R hx_jm = d.F[3][pjm], hz_jm = d.F[5][pjm], hx_km = d.F[3][pkm];
    const R hy_km = d.F[4][pkm], hy_im = d.F[4][pim], hz_im = d.F[5][pim];
    // special-pass cells: no psi update here, own values stored back (the ID
    // is tested after the loads are issued, as in k_lean_e)
    const R E0[3] = {E[0], E[1], E[2]};
    e_cell<R, false, NT>(d, E_BASIC | E_PML, (unsigned)S.b[bi].pml, i, j, k, p, id, hx, hy, hz, hx_jm, hz_jm, hx_km, hy_km, hy_im,
           hz_im, E, true);
    const bool sp = special_id(d, id);
    for (int q = 0; q < 3; q++) E[q] = sp ? E0[q] : E[q];
    st_once<NT>(&d.F[0][p], E[0]);
    st_once<NT>(&d.F[1][p], E[1]);
    st_once<NT>(&d.F[2][p], E[2]);
}

// Batched launches: one launch covers every instance of a batch (gridDim.y
// instances, blockIdx.y = instance); the instance's view is read from the
// device table devs instead of the by-value kernel argument.
template <class R, bool B>
__device__ __forceinline__ const Dev<R> &instance(const Dev<R> &d0, const Dev<R> *devs)
{
    if constexpr (B) return devs[blockIdx.y];
    else return d0;
}

// ---------------------------------------------------------------------------
// Merged half-step kernels: one launch over every tile of the update range.
// A tile inside the interior box L runs the lean stencil; any other tile runs
// the stencil with CPML terms, testing only the axes whose slabs the tile
// reaches (workgroup-uniform mask).  Drude-Lorentz / gain cells are left to
// the special pass (they store their own value back).
//
// Tiles.  Rows are cut into 64-wide tiles (64 x 4 cells, a wavefront per
// x-row) except in the two x-edge columns holding the x CPML slabs, which are
// cut into 32 x 8 tiles: a tile on the CPML path then carries 20 cells
// outside the slab beside the 12 in it, instead of 52 (with 12-cell slabs
// half of all 64-wide tiles would take the CPML path).  The edge columns are
// multiples of 32 cells, so every row segment still starts on a 128-byte
// line in FP32.
//
// Decode.  The launch is 2-D: blockIdx.x is the tile within its plane, read
// from a per-plane table of tile origins (one scalar load per workgroup, no
// integer division), blockIdx.y the plane, blockIdx.z a batch instance.
//
// Measured on the 256^3 FP32 workload (MI355X): 32-wide edge tiles took
// k_full_e 157 -> 134 us and k_full_h 146 -> 133 us (same decode); numbering
// the tiles so that each XCD takes a contiguous z-slab (which helps a
// uniform grid) made both 12-22 us slower here.
// ---------------------------------------------------------------------------
// Waves per SIMD the merged kernels are built for: 8 in FP32, where the
// register allocator keeps them without memory spills (scalar values that do
// not fit the 8-wave budget go to vector-register lanes); FP64 and the
// batched form (whose instance views are read from device memory, where
// every reload is a scalar-cache round trip) left to the allocator.
template <class R, bool B>
constexpr int kFullWaves = sizeof(R) == 4 && !B ? 8 : 1;

// One tile of a plane: its origin, and what the host found about it for the
// half-step the table belongs to (so the kernel branches on one loaded word
// instead of comparing the tile against the slabs and the interior box).
enum : unsigned {
    kTileSlabX = 1u,   // reaches an x CPML slab
    kTileSlabY = 2u,   // reaches a y CPML slab
    kTileInnerXY = 4u  // lies in the interior box's x and y range
};
struct alignas(8) MTileEntry {
    unsigned org;    // x0 | y0 << 15 | (32-wide ? 1 << 31 : 0)
    unsigned flags;  // kTile*
};
struct MGrid {
    const MTileEntry *tiles;  // one plane's tiles
    int per_plane;            // tiles per plane
    int k_lo, planes;         // planes [k_lo, k_lo + planes): H from 0, E from 1
};

struct MTile {
    int x0, y0, k;   // tile origin and plane
    int tw, th;      // 64 x 4 or 32 x 8
    int i, j;        // this thread's cell
    unsigned flags;  // kTile*
};
// tile `tile` of the plane table at plane k (absolute)
__device__ __forceinline__ void mtile_at(const MGrid &g, unsigned tile, int k, MTile &t)
{
    const MTileEntry te = g.tiles[tile];
    const unsigned e = te.org;
    t.flags = te.flags;
    const bool narrow = e >> 31;
    t.x0 = (int)(e & 0x7fffu);
    t.y0 = (int)((e >> 15) & 0xffffu);
    t.k = k;
    t.tw = narrow ? 32 : 64;
    t.th = narrow ? 8 : 4;
    const int lin = threadIdx.x;
    t.i = t.x0 + (narrow ? lin & 31 : lin & 63);
    t.j = t.y0 + (narrow ? lin >> 5 : lin >> 6);
}
// the batch instance of a merged launch (blockIdx.z)
template <class R, bool B>
__device__ __forceinline__ const Dev<R> &instance_z(const Dev<R> &d0, const Dev<R> *devs)
{
    if constexpr (B) return devs[blockIdx.z];
    else return d0;
}

__device__ __forceinline__ bool tile_inside(const Box &L, const MTile &t)
{
    return (t.flags & kTileInnerXY) && t.k >= L.k0 && t.k < L.k1;
}
template <class R>
__device__ __forceinline__ unsigned tile_pml_mask(const Dev<R> &d, const MTile &t, bool efield)
{
    const int o = efield ? 1 : 0;  // E slabs start at index 1 on the low side
    unsigned m = t.flags & (kTileSlabX | kTileSlabY);
    if ((d.npml[2][0] && t.k < d.npml[2][0] + o) || (d.npml[2][1] && t.k >= d.Nz - d.npml[2][1])) m |= 4u;
    return m;
}

// one merged H tile-plane
template <class R, bool CS, bool NT, bool SEL>
__device__ __forceinline__ void h_tile(const Dev<R> &d, const MTile &t, const Box &L, const LeanStretch<R> &c)
{
    const int i = t.i, j = t.j, k = t.k;
    if (i >= d.Nx || j >= d.Ny) return;
    const long long p = cell_index(d.px, d.py, i, j, k);
    const long long pi1 = (d.pbx && i + 1 == d.Nx) ? p + 1 - d.Nx : p + 1;
    const long long pj1 = (d.pby && j + 1 == d.Ny) ? p + d.px - (long long)d.Ny * d.px : p + d.px;
    const long long pk1 = p + d.slab;
    // the ID byte first: loads return in issue order, so the material lookup
    // that depends on it then overlaps the field loads
    const int id = d.id[p];
    const R ex = d.F[0][p], ey = d.F[1][p], ez = d.F[2][p];
    const R ex_j1 = d.F[0][pj1], ez_j1 = d.F[2][pj1], ex_k1 = d.F[0][pk1];
    const R ey_k1 = d.F[1][pk1], ey_i1 = d.F[1][pi1], ez_i1 = d.F[2][pi1];
    R hx = ld_once<NT>(&d.F[3][p]), hy = ld_once<NT>(&d.F[4][p]), hz = ld_once<NT>(&d.F[5][p]);
    if (tile_inside(L, t)) {
        const R aH = coef_aH<SEL>(d, id), bH = coef_bH<SEL>(d, id);
        hx = aH * hx - bH * (ez_j1 - ez) * lean_st<CS>(d, c, S_cHxz, j) + bH * (ey_k1 - ey) * lean_st<CS>(d, c, S_cHxy, k);
        hy = aH * hy - bH * (ex_k1 - ex) * lean_st<CS>(d, c, S_cHyx, k) + bH * (ez_i1 - ez) * lean_st<CS>(d, c, S_cHyz, i);
        hz = aH * hz + bH * (ex_j1 - ex) * lean_st<CS>(d, c, S_cHzx, j) - bH * (ey_i1 - ey) * lean_st<CS>(d, c, S_cHzy, i);
    } else {
        h_cell<R, CS, NT, SEL>(d, true, d.has_pml ? tile_pml_mask(d, t, false) : 0u, i, j, k, p, ex, ey, ez, ex_j1, ez_j1,
                      ex_k1, ey_k1, ey_i1, ez_i1, hx, hy, hz, id, &c);
    }
    st_once<NT>(&d.F[3][p], hx);
    st_once<NT>(&d.F[4][p], hy);
    st_once<NT>(&d.F[5][p], hz);
}

// The H pass runs the planes (and a batch's instances) in descending order,
// the E pass in ascending order: each pass starts where the one before it
// ended, on the planes whose lines the Infinity Cache still holds (256^3:
// FP32 step 0.2806 -> 0.2782 ms, FP64 0.583 -> 0.570 ms).  Cells of one
// pass are independent, so the order changes no value.
template <class R, bool CS, bool NT, bool B = false>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(kFullWaves<R, B>))) void k_full_h(Dev<R> d0, const Dev<R> *devs, MGrid g, Box L, LeanStretch<R> c)
{
    const Dev<R> &d = B ? devs[gridDim.z - 1 - blockIdx.z] : d0;
    MTile t;
    mtile_at(g, blockIdx.x, g.k_lo + (g.planes - 1 - (int)blockIdx.y), t);
    h_tile<R, CS, NT, !B>(d, t, L, c);
}

// one merged E tile-plane
template <class R, bool CS, bool NT, bool SEL>
__device__ __forceinline__ void e_tile(const Dev<R> &d, const MTile &t, const Box &L, const LeanStretch<R> &c)
{
    const int i = t.i, j = t.j, k = t.k;
    if (i < d.ie0 || i >= d.Nx || j < d.je0 || j >= d.Ny || k < 1 || k >= d.Nz) return;
    const long long p = cell_index(d.px, d.py, i, j, k);
    const long long pim = (i == 0) ? p + d.Nx - 1 : p - 1;
    const long long pjm = (j == 0) ? p + (long long)(d.Ny - 1) * d.px : p - d.px;
    const long long pkm = p - d.slab;
    const int id = d.id[p];
    R E[3] = {ld_once<NT>(&d.F[0][p]), ld_once<NT>(&d.F[1][p]), ld_once<NT>(&d.F[2][p])};
    const R hx = d.F[3][p], hy = d.F[4][p], hz = d.F[5][p];
    const R hx_jm = d.F[3][pjm], hz_jm = d.F[5][pjm], hx_km = d.F[3][pkm];
    const R hy_km = d.F[4][pkm], hy_im = d.F[4][pim], hz_im = d.F[5][pim];
    // The special pass's cells keep their values: every cell is computed and
    // theirs are not stored.  (Restoring their old values with a select made
    // the compiler branch around the arithmetic, and the branch's default
    // values waited for the E loads before the stretch constants were
    // fetched.)  The ID test sits in each branch after its loads (hoisted
    // ahead of the branch, it made the CPML tiles wait for the ID byte before
    // issuing their psi loads).
    bool keep;
    if (tile_inside(L, t)) {
        const R aE = coef_aE<SEL>(d, id), bE = coef_bE<SEL>(d, id);
        E[0] = aE * E[0] + bE * (hz - hz_jm) * lean_st<CS>(d, c, S_cExz, j) - bE * (hy - hy_km) * lean_st<CS>(d, c, S_cExy, k);
        E[1] = aE * E[1] + bE * (hx - hx_km) * lean_st<CS>(d, c, S_cEyx, k) - bE * (hz - hz_im) * lean_st<CS>(d, c, S_cEyz, i);
        E[2] = aE * E[2] + bE * (hy - hy_im) * lean_st<CS>(d, c, S_cEzy, i) - bE * (hx - hx_jm) * lean_st<CS>(d, c, S_cEzx, j);
        for (int q = 0; q < 3; q++) E[q] = id == 4 ? (R)0 : E[q];
        keep = special_id(d, id) | (id == 2);  // ID 2 without the ADE module is left as is
    } else {
        e_cell<R, CS, NT, SEL>(d, E_BASIC | E_PML, d.has_pml ? tile_pml_mask(d, t, true) : 0u, i, j, k, p, id, hx, hy, hz,
                      hx_jm, hz_jm, hx_km, hy_km, hy_im, hz_im, E, true, &c);
        keep = special_id(d, id);
    }
    if (keep) return;
    st_once<NT>(&d.F[0][p], E[0]);
    st_once<NT>(&d.F[1][p], E[1]);
    st_once<NT>(&d.F[2][p], E[2]);
}

template <class R, bool CS, bool NT, bool B = false>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(kFullWaves<R, B>))) void k_full_e(Dev<R> d0, const Dev<R> *devs, MGrid g, Box L, LeanStretch<R> c)
{
    const Dev<R> &d = instance_z<R, B>(d0, devs);
    MTile t;
    mtile_at(g, blockIdx.x, g.k_lo + (int)blockIdx.y, t);
    e_tile<R, CS, NT, !B>(d, t, L, c);
}

// Drude-Lorentz and gain cells, all of them: device cell index, with bit
// kSpecialPml set for a cell outside the interior box (its CPML slab terms
// follow the ADE, as in the full update).
constexpr int kSpecialPml = 62;
struct SpecialList {
    const long long *cells;
    int n, pad;
};
// one entry v of the list; ND, NL >= 0: every instance's Drude-Lorentz cells
// have exactly ND Drude and NL Lorentz poles (e_special); -1: any counts
// (e_full)
template <class R, int ND, int NL>
__device__ __forceinline__ void special_cell(const Dev<R> &d, long long v)
{
    const long long p = v & ((1LL << kSpecialPml) - 1);
    const int k = (int)(p / d.slab);
    const long long r = p - (long long)k * d.slab;
    const int j = (int)(r / d.px), i = (int)(r - (long long)j * d.px);
    const unsigned pm = (v >> kSpecialPml) & 1 ? 7u : 0u;
    if constexpr (ND < 0) e_full(d, i, j, k, pm);
    else e_special<R, ND, NL>(d, i, j, k, pm);
}

template <class R, bool B = false, int ND = -1, int NL = -1>
__global__ __launch_bounds__(256) void k_special_e(Dev<R> d0, const Dev<R> *devs, const long long *cells0, int n0,
                                                   const SpecialList *lists)
{
    const Dev<R> &d = instance<R, B>(d0, devs);
    const long long *cells = B ? lists[blockIdx.y].cells : cells0;
    const int n = B ? lists[blockIdx.y].n : n0;
    const int t = blockIdx.x * 256 + threadIdx.x;
    if (t >= n) return;
    special_cell<R, ND, NL>(d, cells[t]);
}

// TF_SF.cpp:966-986
__device__ __forceinline__ double inc_interp(const double *tab, int len, double retarded)
{
    if (retarded <= 0) return 0.0;
    int nt = (int)floor(retarded);
    if (nt > len - 2) nt = len - 2;
    const double resid = retarded - (double)nt;
    return ((1.0 - resid) * tab[nt] + resid * tab[nt + 1]);
}

struct TfTarget {
    long long cell;
    int comp, first, count, pad;
};

// TF/SF face corrections (TF_SF.cpp:988-1089 for E, :1201-1301 for H): one
// thread per corrected field value, its face terms applied in the reference's
// face order.  Incident values are interpolated in FP64.
template <class R>
__device__ __forceinline__ void tfsf_apply(const Dev<R> &d, const TfTarget &g, const double *coef, const double *delay,
                                           const double *prof, const double *tab, int tablen, int n)
{
    R *F = d.F[g.comp];
    R v = F[g.cell];
    for (int q = g.first; q < g.first + g.count; q++) {
        const double inc = inc_interp(tab, tablen, (double)n - delay[q]);
        v = v + (R)(coef[q] * inc * prof[q]);
    }
    F[g.cell] = v;
}
// The FP32 engine's compact form of the same lists: an 8-byte target and
// 12-byte terms (against 24 + 24 bytes), the incident tables in FP32.  A
// term's delay d is split once on the host into the table index of its
// retarded time at step 0 and an interpolation weight, so that the index and
// weight at step n are exact (n - base, w) whatever the magnitude of n:
// retarded = n - d; floor = n - base with base = ceil(d); the weight of entry
// floor + 1 is w = base - d (0 when d is whole); no incident wave while
// retarded <= 0, i.e. floor < 0, or floor == 0 with w == 0.
struct TfTarget32 {
    unsigned cell;  // device index (grids below 2^32 cells)
    unsigned meta;  // component (0..5) << 29 | term count (0..3) << 27 | first term
};
struct TfTerm32 {
    float cp;  // signed C * amplitude * transverse profile
    int base;
    float w;
};
__device__ __forceinline__ void tfsf_apply32(const Dev<float> &d, const TfTarget32 g, const TfTerm32 *terms,
                                             const float *tab, int tablen, int n)
{
    float *F = d.F[g.meta >> 29];
    float v = F[g.cell];
    const int first = (int)(g.meta & ((1u << 27) - 1)), cnt = (int)((g.meta >> 27) & 3u);
    for (int q = first; q < first + cnt; q++) {
        const TfTerm32 t = terms[q];
        int nt = n - t.base;
        if (nt < 0 || (nt == 0 && t.w == 0.0f)) continue;
        if (nt > tablen - 2) nt = tablen - 2;
        v = v + t.cp * ((1.0f - t.w) * tab[nt] + t.w * tab[nt + 1]);
    }
    F[g.cell] = v;
}
template <bool B = false>
__global__ __launch_bounds__(256) void k_tfsf32(Dev<float> d0, const Dev<float> *devs, const TfTarget32 *tg, int ntg,
                                                const TfTerm32 *terms, const float *tab, int tablen, const int *nbase,
                                                int n)
{
    const Dev<float> &d = instance<float, B>(d0, devs);
    if (nbase) n += *nbase;
    const int t = blockIdx.x * 256 + threadIdx.x;
    if (t >= ntg) return;
    tfsf_apply32(d, tg[t], terms, tab, tablen, n);
}

template <class R, bool B = false>
__global__ __launch_bounds__(256) void k_tfsf(Dev<R> d0, const Dev<R> *devs, const TfTarget *tg, int ntg,
                                              const double *coef, const double *delay, const double *prof,
                                              const double *tab, int tablen, const int *nbase, int n)
{
    const Dev<R> &d = instance<R, B>(d0, devs);
    if (nbase) n += *nbase;  // replayed step blocks: step = block start + offset
    const int t = blockIdx.x * 256 + threadIdx.x;
    if (t >= ntg) return;
    tfsf_apply(d, tg[t], coef, delay, prof, tab, tablen, n);
}

// the six field values at a list of cells, [point][6] (Ports.cpp:88-99 gather
// for a caller that reads the current values), written as O (R, or double)
template <class R, class O = R>
__global__ __launch_bounds__(256) void k_gather(Dev<R> d, const long long *cell, int npts, O *out)
{
    const int t = blockIdx.x * 256 + threadIdx.x;
    if (t >= npts) return;
    const long long p = cell[t];
    for (int c = 0; c < 6; c++) out[(long long)t * 6 + c] = (O)d.F[c][p];
}

// Non-finite values in the six fields (the serving loop's watchdog): one
// count per workgroup, added to *count (zeroed by the caller's launch of
// k_set_int).  Reads every field once.
template <class R>
__global__ __launch_bounds__(256) void k_nonfinite(Dev<R> d, long long n, int *count)
{
    __shared__ int red[4];
    int bad = 0;
    for (long long q = (long long)blockIdx.x * 256 + threadIdx.x; q < n; q += (long long)gridDim.x * 256)
        for (int c = 0; c < 6; c++) bad += isfinite(d.F[c][q]) ? 0 : 1;
    for (int off = 32; off > 0; off >>= 1) bad += __shfl_down(bad, off, 64);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = bad;
    __syncthreads();
    if (threadIdx.x == 0) {
        const int t = red[0] + red[1] + red[2] + red[3];
        if (t) atomicAdd(count, t);
    }
}

// slot map of the special cells: aux[cell[s]] = s (the rest of the map holds
// -1 from a fill)
__global__ __launch_bounds__(256) void k_slot_scatter(int *aux, const long long *cell, long long n)
{
    const long long s = (long long)blockIdx.x * 256 + threadIdx.x;
    if (s < n) aux[cell[s]] = (int)s;
}

// one device int: the block start of a replayed step graph, or a counter's
// zero (a kernel, not a memset, so it may sit inside a captured graph)
__global__ void k_set_int(int *dst, int v)
{
    if (threadIdx.x == 0 && blockIdx.x == 0) *dst = v;
}

// Soft point sources (main.cpp:284-288).  Sources are grouped by (cell,
// component); one thread adds its group's waveforms one after another in
// source order, as the reference adds sources, so sources sharing a cell
// neither race nor change the rounding.
struct SourceGroup {
    long long cell;
    int comp, first, count, pad;  // members idx[first .. first + count)
};
template <class R>
__device__ __forceinline__ void source_apply(const Dev<R> &d, const SourceGroup &G, const int *idx, const double *wave,
                                             int stride, int n)
{
    R *F = d.F[G.comp];
    R v = F[G.cell];
    for (int q = G.first; q < G.first + G.count; q++) v = v + (R)wave[(long long)idx[q] * stride + n];
    F[G.cell] = v;
}
template <class R>
__global__ void k_sources(Dev<R> d, const SourceGroup *grp, int ngroups, const int *idx, const double *wave,
                          int stride, const int *nbase, int n)
{
    if (nbase) n += *nbase;
    const int g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= ngroups) return;
    source_apply(d, grp[g], idx, wave, stride, n);
}

// One port of a simulation as the port kernels see it.  Points of all ports
// are handled by one launch; port q owns workgroups [bstart, bstart + nblk).
template <class R>
struct PortDesc {
    const long long *cell;
    R *samples;   // [step][point][6] when recording, else null
    R *batch;     // [slot][6][point] samples awaiting the DFT, null without DFT
    double *dft;  // [f][comp][re,im][point] FP64 accumulators
    double *ipart;// [step][nblk] per-workgroup partial sums of |E|^2
    int npts, nblk, bstart, pad;
};

// DFT accumulators of the points [t0, t0 + n) of a port, from the device
// layout [row][point] (row = frequency * 12 + component * 2 + re/im, np
// points per row) into the host layout [point][row]: 32 x 32 tiles through
// LDS, so that reads and writes are both runs along their own rows.
// Workgroup 32 x 8; grid (ceil(n / 32), ceil(rows / 32)).
__global__ __launch_bounds__(256) void k_dft_transpose(const double *src, long long np, int rows, long long t0, int n,
                                                       double *dst)
{
    __shared__ double tile[32][33];
    const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
    const int r0 = blockIdx.y * 32;
    const long long c0 = (long long)blockIdx.x * 32;
    for (int q = ty; q < 32; q += 8) {
        const int r = r0 + q;
        const long long c = c0 + tx;
        if (r < rows && c < n) tile[q][tx] = src[(long long)r * np + t0 + c];
    }
    __syncthreads();
    for (int q = ty; q < 32; q += 8) {
        const long long c = c0 + q;
        const int r = r0 + tx;
        if (r < rows && c < n) dst[c * rows + r] = tile[tx][q];
    }
}

// Port gather (Ports.cpp:88-99, 342-353) into the time-series record and the
// DFT batch, and the per-step sum of |E|^2 (Ports.cpp:512-522) reduced across
// the workgroup with wavefront shuffles into one partial per workgroup.
// one workgroup (blk) of the port sampling of one simulation's ports; every
// thread of the workgroup calls it (it synchronises the workgroup)
template <class R>
__device__ __forceinline__ void port_block(const Dev<R> &d, const PortDesc<R> *ports, int nports, int blk0, int n,
                                           int slot)
{
    __shared__ double red[4];
    int q = 0;
    while (q + 1 < nports && blk0 >= ports[q + 1].bstart) q++;
    const PortDesc<R> P = ports[q];
    const int blk = blk0 - P.bstart;
    const int t = blk * 256 + threadIdx.x;
    double isq = 0.0;
    if (t < P.npts) {
        const long long p = P.cell[t];
        R v[6];
        for (int c = 0; c < 6; c++) v[c] = d.F[c][p];
        if (P.samples)
            for (int c = 0; c < 6; c++) P.samples[((long long)n * P.npts + t) * 6 + c] = v[c];
        if (slot >= 0)
            for (int c = 0; c < 6; c++) P.batch[((long long)slot * 6 + c) * P.npts + t] = v[c];
        isq = (double)(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
    }
    for (int off = 32; off > 0; off >>= 1) isq += __shfl_down(isq, off, 64);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    if (lane == 0) red[wave] = isq;
    __syncthreads();
    if (threadIdx.x == 0) P.ipart[(long long)n * P.nblk + blk] = ((red[0] + red[1]) + red[2]) + red[3];
}
template <class R, bool B = false>
__global__ __launch_bounds__(256) void k_port(Dev<R> d0, const Dev<R> *devs, const PortDesc<R> *ports, int nports,
                                              const int *nbase, int n, int slot)
{
    if (nbase) n += *nbase;
    const Dev<R> &d = instance<R, B>(d0, devs);
    if (B) ports += (long long)blockIdx.y * nports;  // the instance's port table
    port_block(d, ports, nports, (int)blockIdx.x, n, slot);
}

// The passes after a step's E update in one launch, each on its own range of
// workgroups: Drude-Lorentz / gain cells [0, b1), TF/SF-E corrections
// [b1, b2), point sources [b2, b3), port sampling [b3, end).  Used only when
// the cells whose E values these passes write are distinct from pass to pass
// and the ports sample none of them (the host checks), so that running them
// side by side gives every value the reference's sequence gives.  One launch
// instead of four saves three ramp-ups and tails per step.
template <class R>
struct PostE {
    const long long *special;
    int nspecial, b1;
    const TfTarget *tg;       // FP64 lists, or
    const TfTarget32 *tg32;   // the compact FP32 ones when set
    int ntg, b2;
    const double *coef, *delay, *prof, *tab;
    const TfTerm32 *terms32;
    const float *tab32;
    int tablen;
    const SourceGroup *grp;
    int ngroups, b3;
    const int *sidx;
    const double *swave;
    int sstride;
    const PortDesc<R> *ports;
    int nports, slot;
};
// ND, NL as for special_cell; ND = -2: no special cells (the role compiled
// out, so the other roles keep their register budget)
template <class R, int ND, int NL>
__global__ __launch_bounds__(256) void k_post_e(Dev<R> d, PostE<R> P, const int *nbase, int n)
{
    if (nbase) n += *nbase;  // replayed step blocks: step = block start + offset
    const int b = blockIdx.x;
    if (b < P.b1) {
        // the special cells and the TF/SF-E targets from the top planes
        // down: the E pass before this launch ended there (256^3 FP32:
        // 21.7 -> 21.1 us)
        if constexpr (ND >= -1) {
            const int t = (P.b1 - 1 - b) * 256 + threadIdx.x;
            if (t < P.nspecial) special_cell<R, ND, NL>(d, P.special[t]);
        }
    } else if (b < P.b2) {
        const int t = (P.b2 - 1 - b) * 256 + threadIdx.x;
        if (t >= P.ntg) return;
        if constexpr (std::is_same<R, float>::value) {
            if (P.tg32) {
                tfsf_apply32(d, P.tg32[t], P.terms32, P.tab32, P.tablen, n);
                return;
            }
        }
        tfsf_apply(d, P.tg[t], P.coef, P.delay, P.prof, P.tab, P.tablen, n);
    } else if (b < P.b3) {
        const int g = (b - P.b2) * 256 + threadIdx.x;
        if (g < P.ngroups) source_apply(d, P.grp[g], P.sidx, P.swave, P.sstride, n);
    } else {
        port_block(d, P.ports, P.nports, b - P.b3, n, P.slot);  // the whole workgroup
    }
}

// Running DFT over a batch of nb sampled steps n0, n0+stride, ...: each FP64
// accumulator is read once, gets the batch's terms in step order (the same
// additions, in the same order, as accumulating every sampled step), and is
// written once.  E samples carry phase 2 pi f (n+1) dt, H samples (n+1/2) dt.
// A workgroup takes 64 points, the quarter `quarter` of port block blk0 (the
// port sampling's 256-point blocks): their samples are staged in LDS once,
// and the four wavefronts split the frequencies (lane = point), so the
// samples cross HBM once instead of once per frequency.  Every thread of the
// workgroup calls it (it synchronises).
template <class R>
__device__ __forceinline__ void dft_block(const PortDesc<R> *ports, int nports, int blk0, int quarter, int nb, int n0,
                                          int stride, const double *twE, const double *twH, int nf, int ntot)
{
    __shared__ R xs[16 * 6][64];  // [sampled step * 6 + component][point]; nb <= 16
    int q = 0;
    while (q + 1 < nports && blk0 >= ports[q + 1].bstart) q++;
    const PortDesc<R> P = ports[q];
    const long long np_ = P.npts;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    {
        const int t0 = (blk0 - P.bstart) * 256 + quarter * 64;
        if (t0 >= P.npts) return;  // the same for the whole workgroup
        for (int e = threadIdx.x; e < nb * 6 * 64; e += 256) {
            const int row = e >> 6, pt = e & 63;
            xs[row][pt] = t0 + pt < P.npts ? P.batch[(long long)row * np_ + t0 + pt] : (R)0;
        }
        __syncthreads();
        const int t = t0 + lane;
        if (t < P.npts)
            for (int f = wave; f < nf; f += 4) {
                double *acc = P.dft + (long long)f * 12 * np_ + t;
                double a[12];
                for (int w = 0; w < 12; w++) a[w] = acc[w * np_];
                for (int b = 0; b < nb; b++) {
                    const long long n = n0 + (long long)b * stride;
                    const double *wE = twE + ((long long)f * ntot + n) * 2, *wH = twH + ((long long)f * ntot + n) * 2;
                    for (int c = 0; c < 6; c++) {
                        const double *w = c < 3 ? wE : wH;
                        const double v = (double)xs[b * 6 + c][lane];
                        a[2 * c] = a[2 * c] + v * w[0];
                        a[2 * c + 1] = a[2 * c + 1] - v * w[1];
                    }
                }
                for (int w = 0; w < 12; w++) acc[w * np_] = a[w];
            }
    }
}
template <class R, bool B = false>
__global__ __launch_bounds__(256) void k_dft(const PortDesc<R> *ports, int nports, int nb, const int *nbase, int n0,
                                             int stride, const double *twE, const double *twH, int nf, int ntot)
{
    if (nbase) n0 += *nbase;
    if (B) ports += (long long)blockIdx.y * nports;  // the instance's port table
    dft_block(ports, nports, (int)blockIdx.x >> 2, (int)blockIdx.x & 3, nb, n0, stride, twE, twH, nf, ntot);
}

// ---------------------------------------------------------------------------
// Packed launches: one launch per pass of the step over every member of a
// pack, members of any grid size and physics (the serving loop's requests).
// Each member keeps its own device state; the update kernels find their
// member, plane and tile in a per-workgroup work table, the smaller passes
// their member in per-member workgroup ranges.  A member's time index is its
// step at the start of the launch sequence (nstep[member], device memory)
// plus the offset q of the step within the sequence, so one captured sequence
// serves every quantum; a member whose run has ended skips its workgroups.
// Every cell, target and port point is updated by the arithmetic of the
// single-simulation kernels, so a member's results do not depend on the other
// members (bit for bit).
// ---------------------------------------------------------------------------
template <class R>
struct PackMember {
    Dev<R> d;
    MGrid mh, me;                   // merged update tiles
    Box Lh, Le;                     // interior boxes
    LeanStretch<R> sh, se;          // interior stretch factors (when csH / csE)
    int csH, csE;
    int nt;                         // run length
    int ntfE, ntfH, tablen;         // TF/SF targets per half-step, waveform length
    const TfTarget *tfE, *tfH;
    const double *tcoef, *tdelay, *tprof, *einc, *hinc;
    int tf32;                       // FP32 members: the compact lists below
    const TfTarget32 *tfE32, *tfH32;
    const TfTerm32 *tterm32;
    const float *einc32, *hinc32;
    const long long *special;       // Drude-Lorentz / gain cells (kSpecialPml flag)
    int nspecial;
    int nsgroup;                    // point-source groups
    const SourceGroup *sgroup;
    const int *sidx;
    const double *swave;
    const PortDesc<R> *ports;       // this member's port table
    int nports, dft_stride, nf;     // ports; DFT stride and frequencies (0: no DFT)
    const double *twE, *twH;
};

// the member owning workgroup b: largest s with start[s] <= b (start has n + 1
// ascending entries, start[n] = total workgroups)
__device__ __forceinline__ int pack_member(const int *start, int n, int b)
{
    int lo = 0, hi = n;
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (start[mid] <= b) lo = mid;
        else hi = mid;
    }
    return lo;
}
// work entry of the update kernels: member << 48 | plane << 24 | tile
constexpr int kPackPlaneShift = 24;

template <class R, bool NT>
__global__ __launch_bounds__(256) void k_pack_h(const PackMember<R> *pm, const unsigned long long *work,
                                                const int *nstep, int q)
{
    const unsigned long long w = work[blockIdx.x];
    const int s = (int)(w >> 48);
    const PackMember<R> &P = pm[s];
    if (nstep[s] + q >= P.nt) return;
    MTile t;
    mtile_at(P.mh, (unsigned)(w & 0xffffffu), (int)((w >> kPackPlaneShift) & 0xffffffu), t);
    if (P.csH) h_tile<R, true, NT, false>(P.d, t, P.Lh, P.sh);
    else h_tile<R, false, NT, false>(P.d, t, P.Lh, P.sh);
}

template <class R, bool NT>
__global__ __launch_bounds__(256) void k_pack_e(const PackMember<R> *pm, const unsigned long long *work,
                                                const int *nstep, int q)
{
    const unsigned long long w = work[blockIdx.x];
    const int s = (int)(w >> 48);
    const PackMember<R> &P = pm[s];
    if (nstep[s] + q >= P.nt) return;
    MTile t;
    mtile_at(P.me, (unsigned)(w & 0xffffffu), (int)((w >> kPackPlaneShift) & 0xffffffu), t);
    if (P.csE) e_tile<R, true, NT, false>(P.d, t, P.Le, P.se);
    else e_tile<R, false, NT, false>(P.d, t, P.Le, P.se);
}

// TF/SF corrections of every member (h: the H half-step's)
template <class R>
__global__ __launch_bounds__(256) void k_pack_tfsf(const PackMember<R> *pm, const int *start, int nmem,
                                                   const int *nstep, int q, int h)
{
    const int s = pack_member(start, nmem, blockIdx.x);
    const PackMember<R> &P = pm[s];
    const int n = nstep[s] + q;
    if (n >= P.nt) return;
    const int t = (blockIdx.x - start[s]) * 256 + threadIdx.x;
    if (t >= (h ? P.ntfH : P.ntfE)) return;
    if constexpr (std::is_same<R, float>::value) {
        if (P.tf32) {
            tfsf_apply32(P.d, (h ? P.tfH32 : P.tfE32)[t], P.tterm32, h ? P.einc32 : P.hinc32, P.tablen, n);
            return;
        }
    }
    tfsf_apply(P.d, (h ? P.tfH : P.tfE)[t], P.tcoef, P.tdelay, P.tprof, h ? P.einc : P.hinc, P.tablen, n);
}

// Drude-Lorentz and gain cells of every member (the generic cell update: the
// members' pole counts may differ)
template <class R>
__global__ __launch_bounds__(256) void k_pack_special(const PackMember<R> *pm, const int *start, int nmem,
                                                      const int *nstep, int q)
{
    const int s = pack_member(start, nmem, blockIdx.x);
    const PackMember<R> &P = pm[s];
    if (nstep[s] + q >= P.nt) return;
    const int t = (blockIdx.x - start[s]) * 256 + threadIdx.x;
    if (t >= P.nspecial) return;
    const Dev<R> &d = P.d;
    const long long v = P.special[t];
    const long long p = v & ((1LL << kSpecialPml) - 1);
    const int k = (int)(p / d.slab);
    const long long r = p - (long long)k * d.slab;
    const int j = (int)(r / d.px), i = (int)(r - (long long)j * d.px);
    e_full(d, i, j, k, (v >> kSpecialPml) & 1 ? 7u : 0u);
}

template <class R>
__global__ __launch_bounds__(64) void k_pack_sources(const PackMember<R> *pm, const int *start, int nmem,
                                                     const int *nstep, int q)
{
    const int s = pack_member(start, nmem, blockIdx.x);
    const PackMember<R> &P = pm[s];
    const int n = nstep[s] + q;
    if (n >= P.nt) return;
    const int g = (blockIdx.x - start[s]) * 64 + threadIdx.x;
    if (g >= P.nsgroup) return;
    source_apply(P.d, P.sgroup[g], P.sidx, P.swave, P.nt + 1, n);
}

// port sampling of every member: the DFT staging slot is the member's
// sampled-step count modulo 16 (its flushes come every 16 sampled steps)
template <class R>
__global__ __launch_bounds__(256) void k_pack_port(const PackMember<R> *pm, const int *start, int nmem,
                                                   const int *nstep, int q)
{
    const int s = pack_member(start, nmem, blockIdx.x);
    const PackMember<R> &P = pm[s];
    const int n = nstep[s] + q;
    if (n >= P.nt) return;
    const int slot = (P.nf && n % P.dft_stride == 0) ? (n / P.dft_stride) % 16 : -1;
    port_block(P.d, P.ports, P.nports, (int)blockIdx.x - start[s], n, slot);
}

// DFT flush of every member whose staging is full this step, or whose last
// sampled step of the run this is
template <class R>
__global__ __launch_bounds__(256) void k_pack_dft(const PackMember<R> *pm, const int *start, int nmem,
                                                  const int *nstep, int q)
{
    const int blk = (int)blockIdx.x >> 2;  // four workgroups per port block
    const int s = pack_member(start, nmem, blk);
    const PackMember<R> &P = pm[s];
    const int n = nstep[s] + q;
    if (n >= P.nt || !P.nf || n % P.dft_stride) return;
    const int slot = (n / P.dft_stride) % 16;
    const int last = (P.nt - 1) / P.dft_stride * P.dft_stride;
    if (slot != 15 && n != last) return;
    dft_block(P.ports, P.nports, blk - start[s], (int)blockIdx.x & 3, slot + 1, n - slot * P.dft_stride, P.dft_stride,
              P.twE, P.twH, P.nf, P.nt);
}

// every member's step base moves on by the length of the launched sequence
__global__ void k_pack_advance(int *nstep, int nmem, int count)
{
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s < nmem) nstep[s] += count;
}

}  // namespace cfd
