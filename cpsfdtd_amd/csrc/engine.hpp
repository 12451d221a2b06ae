// engine.hpp -- device-side view of one simulation and the kernel interface.
//
// Data layout in HBM (all per simulation):
//   fields Ex..Hz     structure of arrays, x fastest: p = i + px*(j + py*k),
//                     px = Nx+1 rounded up to 32 elements (128-byte rows for
//                     both FP32 and FP64), py = Ny+1; (Nx+1)..px-1 is padding
//   material ID       uint8 in the same layout
//   aux index         int32 in the same layout, slot of a Drude-Lorentz or gain
//                     cell in its state arrays (only when those modules exist)
//   CPML psi          slab boxes, x fastest; thickness n along the slab axis
//   coefficient and stretch tables, TF/SF target lists, port point lists
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>

namespace cfd {

constexpr int kMaxPoles = 8;
constexpr int TX = 64;  // tile width along x: one wavefront
constexpr int TY = 4;   // tile height along y: wavefronts per workgroup
constexpr int kSelMat = 8;  // media whose coefficients the kernels select without a table load

// parts of the E half-step a kernel performs (fused step = all of them)
enum : unsigned {
    E_BASIC = 1u,    // BasicFDTD.cpp:25-45
    E_DL = 2u,       // dispersive cells (ID 2)
    E_GAINPOL = 4u,  // gain polarisation P^{n+1} from E^n (ID 5)
    E_GAINE = 8u,    // E -= (bE/dt) dP
    E_GAINPOP = 16u, // four-level populations
    E_PML = 32u,     // CPML.cpp:529-607
    E_ALL = 63u,
};
enum : unsigned { H_BASIC = 1u, H_PML = 2u, H_ALL = 3u };

// two values of one cell loaded or stored as one aligned 2-word access
template <class R>
struct alignas(2 * sizeof(R)) Pair {
    R a, b;
};

template <class R>
struct Dev {
    int Nx, Ny, Nz, px, py;
    long long slab;  // px*py
    int pbx, pby, ie0, je0;
    R *F[6];
    const uint8_t *id;
    const int *aux;
    const R *aE, *bE, *aH, *bH;
    // with at most kSelMat media (nsel != 0) the same coefficients are also
    // held here, in the kernel argument, and selected by the ID's bits: no
    // table load that has to wait for the ID byte
    int nsel;
    R selE[2][kSelMat], selH[2][kSelMat];  // [a, b][ID]
    const R *st[12];
    // CPML
    int has_pml;
    int npml[3][2];
    // CPML slabs per [E/H][axis]: the low side's grading coefficient pairs
    // (b, c) then the high side's in one array (high side from index pofs),
    // and one psi array of pairs (the two psi of the axis at a cell, loaded
    // and stored as one 2-word access) holding the low slab's box then the
    // high slab's (from qofs), so a cell's side selects an offset from a
    // uniform base, not a pointer
    const Pair<R> *pbc[2][3];
    Pair<R> *q[2][3];                    // [E/H][axis]
    int pofs[2][3];
    long long qofs[2][3];
    R cpmlE, cpmlH;
    int hsign_fixed;
    // Drude-Lorentz
    int nD, nL, ndl;
    R C1, C2, C3, dt;
    R C4[kMaxPoles], C5[kMaxPoles], C6[kMaxPoles], kap[kMaxPoles], beta[kMaxPoles];
    R alpha[kMaxPoles], xi[kMaxPoles], gam[kMaxPoles];
    R *dlst;  // [(var*3 + comp) * ndl + slot]
    // gain (populations FP64)
    int ngain;
    R CPa1, CPa2;
    double g_CPa3;  // P coupling to the inversion, FP64 (subnormal in FP32)
    double g_dt, g_hw, g_Ntot, g_h32, g_h21, g_h10, g_dtWp;
    R *gP, *gP1, *gEprev;  // [comp * ngain + slot]
    double *gN;            // [level * ngain + slot], N0..N3
};

}  // namespace cfd
