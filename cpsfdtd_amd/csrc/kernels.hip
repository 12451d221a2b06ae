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

#include "engine.hpp"

namespace cfd {

enum { S_cExy, S_cEyx, S_cEzx, S_cExz, S_cEzy, S_cEyz, S_cHxy, S_cHyx, S_cHzx, S_cHxz, S_cHzy, S_cHyz };

__device__ __forceinline__ long long cell_index(int px, int py, int i, int j, int k)
{
    return (long long)i + (long long)px * ((long long)j + (long long)py * k);
}

// H half-step: BasicFDTD.cpp:67-81 over [0,N)^3, then CPML.cpp:609-689 in the
// slab tiles (x, then y, then z contributions, as the reference orders them).
template <class R>
__global__ __launch_bounds__(TX *TY) void k_update_h(Dev<R> d, unsigned parts)
{
    const int x0 = blockIdx.x * TX, y0 = blockIdx.y * TY, k = blockIdx.z;
    const int i = x0 + threadIdx.x, j = y0 + threadIdx.y;
    bool tile_pml = false;
    if ((parts & H_PML) && d.has_pml) {
        tile_pml = (d.npml[0][0] && x0 < d.npml[0][0]) || (d.npml[0][1] && x0 + TX > d.Nx - d.npml[0][1]) ||
                   (d.npml[1][0] && y0 < d.npml[1][0]) || (d.npml[1][1] && y0 + TY > d.Ny - d.npml[1][1]) ||
                   (k < d.npml[2][0]) || (d.npml[2][1] && k >= d.Nz - d.npml[2][1]);
    }
    if (!(parts & H_BASIC) && !tile_pml) return;
    if (i >= d.Nx || j >= d.Ny) return;
    const long long p = cell_index(d.px, d.py, i, j, k);
    const long long pi1 = (d.pbx && i + 1 == d.Nx) ? p + 1 - d.Nx : p + 1;
    const long long pj1 = (d.pby && j + 1 == d.Ny) ? p + d.px - (long long)d.Ny * d.px : p + d.px;
    const long long pk1 = p + d.slab;
    const R ex = d.F[0][p], ey = d.F[1][p], ez = d.F[2][p];
    R hx = d.F[3][p], hy = d.F[4][p], hz = d.F[5][p];
    if (parts & H_BASIC) {
        const int id = d.id[p];
        const R aH = d.aH[id], bH = d.bH[id];
        hx = aH * hx - bH * (d.F[2][pj1] - ez) * d.st[S_cHxz][j] + bH * (d.F[1][pk1] - ey) * d.st[S_cHxy][k];
        hy = aH * hy - bH * (d.F[0][pk1] - ex) * d.st[S_cHyx][k] + bH * (d.F[2][pi1] - ez) * d.st[S_cHyz][i];
        hz = aH * hz + bH * (d.F[0][pj1] - ex) * d.st[S_cHzx][j] - bH * (d.F[1][pi1] - ey) * d.st[S_cHzy][i];
    }
    if (tile_pml) {
        const R cH = d.cpmlH;
        const bool fixed = d.hsign_fixed;
        for (int e = 0; e < 2; e++) {  // x slabs: Hy += psi(dEz/dx), Hz -= psi(dEy/dx)
            const int n = d.npml[0][e];
            const int l = e == 0 ? i : i - (d.Nx - n);
            if (n && l >= 0 && l < n) {
                const R b = d.pb[1][0][e][l], c = d.pc[1][0][e][l];
                const long long qi = l + (long long)n * (j + (long long)d.py * k);
                R *q1 = d.q[1][0][e][0], *q2 = d.q[1][0][e][1];
                const R a1 = b * q1[qi] + c * (d.F[2][pi1] - ez);
                q1[qi] = a1;
                hy = hy + a1 * cH;
                const R a2 = b * q2[qi] + c * (d.F[1][pi1] - ey);
                q2[qi] = a2;
                hz = hz - a2 * cH;
            }
        }
        for (int e = 0; e < 2; e++) {  // y slabs: Hz -/+ psi(dEx/dy), Hx +/- psi(dEz/dy)
            const int n = d.npml[1][e];
            const int l = e == 0 ? j : j - (d.Ny - n);
            if (n && l >= 0 && l < n) {
                const R b = d.pb[1][1][e][l], c = d.pc[1][1][e][l];
                const long long qi = i + (long long)d.px * (l + (long long)n * k);
                R *q1 = d.q[1][1][e][0], *q2 = d.q[1][1][e][1];
                const R a1 = b * q1[qi] + c * (d.F[0][pj1] - ex);
                q1[qi] = a1;
                hz = fixed ? hz + a1 * cH : hz - a1 * cH;
                const R a2 = b * q2[qi] + c * (d.F[2][pj1] - ez);
                q2[qi] = a2;
                hx = fixed ? hx - a2 * cH : hx + a2 * cH;
            }
        }
        for (int e = 0; e < 2; e++) {  // z slabs: Hx -/+ psi(dEy/dz), Hy +/- psi(dEx/dz)
            const int n = d.npml[2][e];
            const int l = e == 0 ? k : k - (d.Nz - n);
            if (n && l >= 0 && l < n) {
                const R b = d.pb[1][2][e][l], c = d.pc[1][2][e][l];
                const long long qi = i + (long long)d.px * (j + (long long)d.py * l);
                R *q1 = d.q[1][2][e][0], *q2 = d.q[1][2][e][1];
                const R a1 = b * q1[qi] + c * (d.F[1][pk1] - ey);
                q1[qi] = a1;
                hx = fixed ? hx + a1 * cH : hx - a1 * cH;
                const R a2 = b * q2[qi] + c * (d.F[0][pk1] - ex);
                q2[qi] = a2;
                hy = fixed ? hy - a2 * cH : hy + a2 * cH;
            }
        }
    }
    d.F[3][p] = hx;
    d.F[4][p] = hy;
    d.F[5][p] = hz;
}

// E half-step over [1,N)^3 (x/y from 0 when periodic): four-level gain
// polarisation from E^n, BasicFDTD.cpp:25-45 (ID 4 zeroed, ID 2 left to the
// Drude-Lorentz ADE), the ADE for ID 2, the gain current and populations for
// ID 5, then CPML.cpp:529-607 in the slab tiles.
template <class R>
__global__ __launch_bounds__(TX *TY) void k_update_e(Dev<R> d, unsigned parts)
{
    const int x0 = blockIdx.x * TX, y0 = blockIdx.y * TY, k = blockIdx.z;
    const int i = x0 + threadIdx.x, j = y0 + threadIdx.y;
    bool tile_pml = false;
    if ((parts & E_PML) && d.has_pml) {
        tile_pml = (d.npml[0][0] && x0 <= d.npml[0][0]) || (d.npml[0][1] && x0 + TX > d.Nx - d.npml[0][1]) ||
                   (d.npml[1][0] && y0 <= d.npml[1][0]) || (d.npml[1][1] && y0 + TY > d.Ny - d.npml[1][1]) ||
                   (k <= d.npml[2][0]) || (d.npml[2][1] && k >= d.Nz - d.npml[2][1]);
    }
    if (!(parts & ~E_PML) && !tile_pml) return;
    if (i < d.ie0 || i >= d.Nx || j < d.je0 || j >= d.Ny || k < 1 || k >= d.Nz) return;
    const long long p = cell_index(d.px, d.py, i, j, k);
    const int id = d.id[p];
    const unsigned gparts = parts & (E_GAINPOL | E_GAINE | E_GAINPOP);
    const bool gain = gparts && d.ngain && id == 5;
    const bool dl = (parts & E_DL) && d.ndl && id == 2;
    if (!(parts & (E_BASIC | E_PML)) && !gain && !dl) return;
    const long long pim = (i == 0) ? p + d.Nx - 1 : p - 1;
    const long long pjm = (j == 0) ? p + (long long)(d.Ny - 1) * d.px : p - d.px;
    const long long pkm = p - d.slab;
    R E[3] = {d.F[0][p], d.F[1][p], d.F[2][p]};
    const R hx = d.F[3][p], hy = d.F[4][p], hz = d.F[5][p];
    R En[3] = {E[0], E[1], E[2]};  // E^n, kept for the gain populations
    int gs = -1;
    if (gain) {
        gs = d.aux[p];
        if (parts & E_GAINPOL) {
            const double dN = d.gN[1LL * d.ngain + gs] - d.gN[2LL * d.ngain + gs];
            const R dNr = (R)dN;
            for (int c = 0; c < 3; c++) {
                const long long o = (long long)c * d.ngain + gs;
                const R Pn = d.gP[o];
                d.gP[o] = d.CPa1 * Pn + d.CPa2 * d.gP1[o] + d.CPa3 * dNr * E[c];
                d.gP1[o] = Pn;
                if (!(parts & E_GAINPOP)) d.gEprev[o] = E[c];
            }
        }
    }
    if (parts & E_BASIC) {
        if (id == 4) {
            E[0] = 0; E[1] = 0; E[2] = 0;
        } else if (id != 2) {
            const R aE = d.aE[id], bE = d.bE[id];
            E[0] = aE * E[0] + bE * (hz - d.F[5][pjm]) * d.st[S_cExz][j] - bE * (hy - d.F[4][pkm]) * d.st[S_cExy][k];
            E[1] = aE * E[1] + bE * (hx - d.F[3][pkm]) * d.st[S_cEyx][k] - bE * (hz - d.F[5][pim]) * d.st[S_cEyz][i];
            E[2] = aE * E[2] + bE * (hy - d.F[4][pim]) * d.st[S_cEzy][i] - bE * (hx - d.F[3][pjm]) * d.st[S_cEzx][j];
        }
    }
    if (dl) {  // Drude_Lorentz.cpp:172-206, all three components, C1/C5 per F9/F10
        const int s = d.aux[p];
        const R dA[3] = {hz - d.F[5][pjm], hx - d.F[3][pkm], hy - d.F[4][pim]};
        const R cA[3] = {d.st[S_cExz][j], d.st[S_cEyx][k], d.st[S_cEzy][i]};
        const R dB[3] = {hy - d.F[4][pkm], hz - d.F[5][pim], hx - d.F[3][pjm]};
        const R cB[3] = {d.st[S_cExy][k], d.st[S_cEyz][i], d.st[S_cEzx][j]};
        const int nD = d.nD, nL = d.nL;
        const long long ndl = d.ndl;
        for (int c = 0; c < 3; c++) {
            R *JD = d.dlst + (0LL * 3 + c) * ndl + s;               // var v at JD + v*3*ndl
            R *JL = d.dlst + ((long long)nD * 3 + c) * ndl + s;
            R *JL1 = d.dlst + ((long long)(nD + nL) * 3 + c) * ndl + s;
            R *E1 = d.dlst + ((long long)(nD + 2 * nL) * 3 + c) * ndl + s;
            const long long vs = 3 * ndl;
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
        const R coef = d.bE[id] / d.dt;
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
    if (tile_pml) {
        const R cE = d.cpmlE;
        for (int e = 0; e < 2; e++) {  // x slabs: Ey -= psi(dHz/dx), Ez += psi(dHy/dx)
            const int n = d.npml[0][e];
            const int l = e == 0 ? i - 1 : i - (d.Nx - n);
            if (n && l >= 0 && l < n) {
                const int ci = e == 0 ? i : l;
                const R b = d.pb[0][0][e][ci], c = d.pc[0][0][e][ci];
                const long long qi = l + (long long)n * (j + (long long)d.py * k);
                R *q1 = d.q[0][0][e][0], *q2 = d.q[0][0][e][1];
                const R a1 = b * q1[qi] + c * (hz - d.F[5][pim]);
                q1[qi] = a1;
                E[1] = E[1] - a1 * cE;
                const R a2 = b * q2[qi] + c * (hy - d.F[4][pim]);
                q2[qi] = a2;
                E[2] = E[2] + a2 * cE;
            }
        }
        for (int e = 0; e < 2; e++) {  // y slabs: Ez -= psi(dHx/dy), Ex += psi(dHz/dy)
            const int n = d.npml[1][e];
            const int l = e == 0 ? j - 1 : j - (d.Ny - n);
            if (n && l >= 0 && l < n) {
                const int cj = e == 0 ? j : l;
                const R b = d.pb[0][1][e][cj], c = d.pc[0][1][e][cj];
                const long long qi = i + (long long)d.px * (l + (long long)n * k);
                R *q1 = d.q[0][1][e][0], *q2 = d.q[0][1][e][1];
                const R a1 = b * q1[qi] + c * (hx - d.F[3][pjm]);
                q1[qi] = a1;
                E[2] = E[2] - a1 * cE;
                const R a2 = b * q2[qi] + c * (hz - d.F[5][pjm]);
                q2[qi] = a2;
                E[0] = E[0] + a2 * cE;
            }
        }
        for (int e = 0; e < 2; e++) {  // z slabs: Ex -= psi(dHy/dz), Ey += psi(dHx/dz)
            const int n = d.npml[2][e];
            const int l = e == 0 ? k - 1 : k - (d.Nz - n);
            if (n && l >= 0 && l < n) {
                const int ck = e == 0 ? k : l;
                const R b = d.pb[0][2][e][ck], c = d.pc[0][2][e][ck];
                const long long qi = i + (long long)d.px * (j + (long long)d.py * l);
                R *q1 = d.q[0][2][e][0], *q2 = d.q[0][2][e][1];
                const R a1 = b * q1[qi] + c * (hy - d.F[4][pkm]);
                q1[qi] = a1;
                E[0] = E[0] - a1 * cE;
                const R a2 = b * q2[qi] + c * (hx - d.F[3][pkm]);
                q2[qi] = a2;
                E[1] = E[1] + a2 * cE;
            }
        }
    }
    d.F[0][p] = E[0];
    d.F[1][p] = E[1];
    d.F[2][p] = E[2];
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
__global__ __launch_bounds__(256) void k_tfsf(Dev<R> d, const TfTarget *tg, int ntg, const double *coef,
                                              const double *delay, const double *prof, const double *tab, int tablen,
                                              int n)
{
    const int t = blockIdx.x * 256 + threadIdx.x;
    if (t >= ntg) return;
    const TfTarget g = tg[t];
    R *F = d.F[g.comp];
    R v = F[g.cell];
    for (int q = g.first; q < g.first + g.count; q++) {
        const double inc = inc_interp(tab, tablen, (double)n - delay[q]);
        v = v + (R)(coef[q] * inc * prof[q]);
    }
    F[g.cell] = v;
}

// soft point sources (main.cpp:284-288)
template <class R>
__global__ void k_sources(Dev<R> d, const long long *cell, const int *comp, const double *wave, int nsrc, int stride, int n)
{
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= nsrc) return;
    R *F = d.F[comp[q]];
    F[cell[q]] = F[cell[q]] + (R)wave[(long long)q * stride + n];
}

// Port gather (Ports.cpp:88-99, 342-353), running FP64 DFT and the per-step
// sum of |E|^2 (Ports.cpp:512-522), reduced across the workgroup with
// wavefront shuffles into one partial per workgroup (fixed order).
template <class R>
__global__ __launch_bounds__(256) void k_port(Dev<R> d, const long long *cell, int npts, int n, R *samples,
                                              double *dft, const double *twE, const double *twH, int nf, int ntot,
                                              int do_dft, double *ipart)
{
    __shared__ double red[4];
    const int t = blockIdx.x * 256 + threadIdx.x;
    double isq = 0.0;
    if (t < npts) {
        const long long p = cell[t];
        R v[6];
        for (int c = 0; c < 6; c++) v[c] = d.F[c][p];
        if (samples)
            for (int c = 0; c < 6; c++) samples[((long long)n * npts + t) * 6 + c] = v[c];
        isq = (double)(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
        if (do_dft) {  // accumulators [f][comp][re,im][point]: consecutive lanes, consecutive words
            const long long np_ = npts;
            for (int f = 0; f < nf; f++) {
                const double *wE = twE + ((long long)f * ntot + n) * 2, *wH = twH + ((long long)f * ntot + n) * 2;
                double *acc = dft + (long long)f * 12 * np_ + t;
                for (int c = 0; c < 6; c++) {
                    const double *w = c < 3 ? wE : wH;
                    double *re = acc + (2 * c) * np_, *im = acc + (2 * c + 1) * np_;
                    *re = *re + (double)v[c] * w[0];
                    *im = *im - (double)v[c] * w[1];
                }
            }
        }
    }
    for (int off = 32; off > 0; off >>= 1) isq += __shfl_down(isq, off, 64);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    if (lane == 0) red[wave] = isq;
    __syncthreads();
    if (threadIdx.x == 0) ipart[(long long)n * gridDim.x + blockIdx.x] = ((red[0] + red[1]) + red[2]) + red[3];
}

}  // namespace cfd
