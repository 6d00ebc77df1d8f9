// FDTD field update on the periodic box.
//
// UpdateBHalfFunctor (B -= curl E * 0.5 * dt) and UpdateEFunctor
// (E += curl B * c^2 * dt) of fields/MaxwellSolver/FDTD/FDTDBase.kernel:50-124,
// with the Yee curls (Curl<Forward> for E, Curl<Backward> for B;
// fields/differentiation/Curl.hpp:89, ForwardDerivative.hpp:44-65,
// BackwardDerivative.hpp:44-65) or the Lehe curl of E
// (fields/MaxwellSolver/Lehe/Derivative.hpp:65-228), and AddCurrentDensity
// with currentInterpolation::None (E += coeff * J, None.hpp:56-65).
// Fields are SoA planes, x fastest; neighbours wrap periodically, except
// beyond the border of an open axis, where every field reads zero (the
// reference's guard cells of a non-periodic border are never written).
#include "pic_kernels.hpp"

namespace pic {

struct Idx {
    GridDev g;
    int x, y, z;
    __device__ __forceinline__ size_t at(int dx, int dy, int dz) const {
        int xx = x + dx, yy = y + dy, zz = z + dz;
        xx = xx < 0 ? xx + g.nx : (xx >= g.nx ? xx - g.nx : xx);
        yy = yy < 0 ? yy + g.ny : (yy >= g.ny ? yy - g.ny : yy);
        zz = zz < 0 ? zz + g.nz : (zz >= g.nz ? zz - g.nz : zz);
        return (size_t)xx + (size_t)g.nx * ((size_t)yy + (size_t)g.ny * zz);
    }
    // value of f at the offset cell; zero beyond an open border
    __device__ __forceinline__ float get(const float* f, int dx, int dy, int dz) const {
        return outsideOpen(g, x + dx, y + dy, z + dz) ? 0.0f : f[at(dx, dy, dz)];
    }
};

__device__ __forceinline__ void unit(int d, int& ux, int& uy, int& uz) {
    ux = d == 0;
    uy = d == 1;
    uz = d == 2;
}

// forward difference of component f along axis d at offset (ox,oy,oz)
__device__ __forceinline__ float fwd(const float* f, const Idx& I, int d, int ox, int oy, int oz, float step) {
    int ux, uy, uz;
    unit(d, ux, uy, uz);
    return (I.get(f, ox + ux, oy + uy, oz + uz) - I.get(f, ox, oy, oz)) / step;
}

// Lehe derivative along axis d of component f (Lehe/Derivative.hpp:132-145 for
// d == Cherenkov-free axis, :212-224 otherwise)
__device__ __forceinline__ float leheDeriv(const float* f, const Idx& I, int d, const FieldArgs& a) {
    int const L = a.lehe_dir;
    float const step = a.cell[d];
    if (d == L) {
        int const d1 = (d + 1) % 3, d2 = (d + 2) % 3;
        int u1x, u1y, u1z, u2x, u2y, u2z, u0x, u0y, u0z;
        unit(d1, u1x, u1y, u1z);
        unit(d2, u2x, u2y, u2z);
        unit(d, u0x, u0y, u0z);
        float r = a.lehe_alpha * fwd(f, I, d, 0, 0, 0, step);
        r = r + a.lehe_beta1 * fwd(f, I, d, u1x, u1y, u1z, step);
        r = r + a.lehe_beta1 * fwd(f, I, d, -u1x, -u1y, -u1z, step);
        r = r + a.lehe_beta2 * fwd(f, I, d, u2x, u2y, u2z, step);
        r = r + a.lehe_beta2 * fwd(f, I, d, -u2x, -u2y, -u2z, step);
        r = r + a.lehe_delta * (I.get(f, 2 * u0x, 2 * u0y, 2 * u0z) - I.get(f, -u0x, -u0y, -u0z)) / step;
        return r;
    }
    int ulx, uly, ulz;
    unit(L, ulx, uly, ulz);
    float const beta = 0.125f;
    float const alpha = 1.0f - 2.0f * beta;
    float r = alpha * fwd(f, I, d, 0, 0, 0, step);
    r = r + beta * fwd(f, I, d, ulx, uly, ulz, step);
    r = r + beta * fwd(f, I, d, -ulx, -uly, -ulz, step);
    return r;
}

template <int SOLVER>
__device__ __forceinline__ float derivE(const float* f, const Idx& I, int d, const FieldArgs& a) {
    if (SOLVER == PIC_SOLVER_LEHE) return leheDeriv(f, I, d, a);
    return fwd(f, I, d, 0, 0, 0, a.cell[d]);
}

// a cell known to lie outside every PML layer (FieldArgs::inner)
__device__ __forceinline__ bool innerCell(const FieldArgs& a, int x, int y, int z) {
    return a.inner && x >= a.in0[0] && x < a.in1[0] && y >= a.in0[1] && y < a.in1[1] && z >= a.in0[2] &&
           z < a.in1[2];
}

template <int SOLVER>
__global__ __launch_bounds__(256) void kBHalf(FieldArgs a) {
    if (halted(a.g)) return;
    size_t const n = (size_t)a.g.nx * a.g.ny * a.g.nz;
    size_t const i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    Idx I{a.g, (int)(i % a.g.nx), (int)((i / a.g.nx) % a.g.ny), (int)(i / ((size_t)a.g.nx * a.g.ny))};
    // curl = (dFdy.z - dFdz.y, dFdz.x - dFdx.z, dFdx.y - dFdy.x)
    float const cx = derivE<SOLVER>(a.E[2], I, 1, a) - derivE<SOLVER>(a.E[1], I, 2, a);
    float const cy = derivE<SOLVER>(a.E[0], I, 2, a) - derivE<SOLVER>(a.E[2], I, 0, a);
    float const cz = derivE<SOLVER>(a.E[1], I, 0, a) - derivE<SOLVER>(a.E[0], I, 1, a);
    a.B[0][i] = a.B[0][i] - cx * 0.5f * a.dt;
    a.B[1][i] = a.B[1][i] - cy * 0.5f * a.dt;
    a.B[2][i] = a.B[2][i] - cz * 0.5f * a.dt;
}

__device__ __forceinline__ float bwd(const float* f, const Idx& I, int d, float step) {
    int ux, uy, uz;
    unit(d, ux, uy, uz);
    return (I.get(f, 0, 0, 0) - I.get(f, -ux, -uy, -uz)) / step;
}

__global__ __launch_bounds__(256) void kEFull(FieldArgs a) {
    if (halted(a.g)) return;
    size_t const n = (size_t)a.g.nx * a.g.ny * a.g.nz;
    size_t const i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    Idx I{a.g, (int)(i % a.g.nx), (int)((i / a.g.nx) % a.g.ny), (int)(i / ((size_t)a.g.nx * a.g.ny))};
    float const cx = bwd(a.B[2], I, 1, a.cell[1]) - bwd(a.B[1], I, 2, a.cell[2]);
    float const cy = bwd(a.B[0], I, 2, a.cell[2]) - bwd(a.B[2], I, 0, a.cell[0]);
    float const cz = bwd(a.B[1], I, 0, a.cell[0]) - bwd(a.B[0], I, 1, a.cell[1]);
    a.E[0][i] = a.E[0][i] + cx * a.c2 * a.dt;
    a.E[1][i] = a.E[1][i] + cy * a.c2 * a.dt;
    a.E[2][i] = a.E[2][i] + cz * a.c2 * a.dt;
}

// ---------------------------------------------------------------------------
// PML (absorber/pml/Pml.kernel:385-590). A cell lies in the layer when the
// product of its b coefficients along x, y, z is not 1 (isInPML); there the
// convolution fields psi advance (b psi + c dF) and the curl's derivatives
// are divided by kappa, in UpdateEFunctor's and UpdateBHalfFunctor's order of
// operations. Outside it the cell takes kEFull's / kBHalf's arithmetic.
// ---------------------------------------------------------------------------
struct PmlCoeffs {
    float b[3], c[3], kappa[3];
};

// f / kappa: a correctly rounded division by 1 returns f, so the default
// kappa of 1 skips the division with the same result
__device__ __forceinline__ float divKappa(float f, float kappa) { return kappa == 1.0f ? f : f / kappa; }

__device__ __forceinline__ PmlCoeffs pmlCoeffs(const float* __restrict__ tab, const GridDev& g, int x, int y, int z) {
    int const total = g.nx + g.ny + g.nz;
    int const at[3] = {x, g.nx + y, g.nx + g.ny + z};
    PmlCoeffs k;
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        k.b[d] = tab[at[d]];
        k.c[d] = tab[total + at[d]];
        k.kappa[d] = tab[2 * total + at[d]];
    }
    return k;
}

// B half step of cell i from the derivatives of E along each axis (dEdxY:
// of E.y along x, ...): the PML's update inside the layer, kBHalf's
// (B -= curl E * 0.5 * dt) outside it
__device__ __forceinline__ void bHalfPmlApply(const FieldArgs& a, const PmlArgs& p, int updatePsi, int x, int y, int z,
                                              size_t i, float dEdxY, float dEdxZ, float dEdyX, float dEdyZ,
                                              float dEdzX, float dEdzY) {
    PmlCoeffs const k = pmlCoeffs(p.tabB, a.g, x, y, z);
    if (k.b[0] * k.b[1] * k.b[2] == 1.0f) {  // outside the layer: kBHalf
        float const cx = dEdyZ - dEdzY;
        float const cy = dEdzX - dEdxZ;
        float const cz = dEdxY - dEdyX;
        a.B[0][i] = a.B[0][i] - cx * 0.5f * a.dt;
        a.B[1][i] = a.B[1][i] - cy * 0.5f * a.dt;
        a.B[2][i] = a.B[2][i] - cz * 0.5f * a.dt;
        return;
    }
    float yx = p.psiB[0][i], zx = p.psiB[1][i], xy = p.psiB[2][i];
    float zy = p.psiB[3][i], xz = p.psiB[4][i], yz = p.psiB[5][i];
    if (updatePsi) {
        yx = k.b[0] * yx + k.c[0] * dEdxZ;
        zx = k.b[0] * zx + k.c[0] * dEdxY;
        xy = k.b[1] * xy + k.c[1] * dEdyZ;
        zy = k.b[1] * zy + k.c[1] * dEdyX;
        xz = k.b[2] * xz + k.c[2] * dEdzY;
        yz = k.b[2] * yz + k.c[2] * dEdzX;
        p.psiB[0][i] = yx;
        p.psiB[1][i] = zx;
        p.psiB[2][i] = xy;
        p.psiB[3][i] = zy;
        p.psiB[4][i] = xz;
        p.psiB[5][i] = yz;
    }
    float const halfDt = 0.5f * a.dt;
    a.B[0][i] = a.B[0][i] + halfDt * (divKappa(dEdzY, k.kappa[2]) - divKappa(dEdyZ, k.kappa[1]) + xz - xy);
    a.B[1][i] = a.B[1][i] + halfDt * (divKappa(dEdxZ, k.kappa[0]) - divKappa(dEdzX, k.kappa[2]) + yx - yz);
    a.B[2][i] = a.B[2][i] + halfDt * (divKappa(dEdyX, k.kappa[1]) - divKappa(dEdxY, k.kappa[0]) + zy - zx);
}

template <int SOLVER>
__device__ __forceinline__ void bHalfPmlAt(const FieldArgs& a, const PmlArgs& p, int updatePsi, int x, int y, int z) {
    Idx I{a.g, x, y, z};
    size_t const i = (size_t)x + (size_t)a.g.nx * ((size_t)y + (size_t)a.g.ny * (size_t)z);
    // dEdx = curl.xDerivative(E) etc.: the derivative of every component along one axis
    bHalfPmlApply(a, p, updatePsi, x, y, z, i, derivE<SOLVER>(a.E[1], I, 0, a), derivE<SOLVER>(a.E[2], I, 0, a),
                  derivE<SOLVER>(a.E[0], I, 1, a), derivE<SOLVER>(a.E[2], I, 1, a), derivE<SOLVER>(a.E[0], I, 2, a),
                  derivE<SOLVER>(a.E[1], I, 2, a));
}

// E full step with PML of cell i into Eout (a.E itself for the stage call)
// from the backward derivatives of B along each axis (dBdxY: of B.y along
// x, ...): the PML's update inside the layer, kEFull's outside it; ADDJ:
// then J -> E as kAddCurrent (None), in the same pass
template <bool ADDJ>
__device__ __forceinline__ void eFullPmlApply(const FieldArgs& a, const PmlArgs& p, float* const* Eout, int x, int y,
                                              int z, size_t i, float dBdxY, float dBdxZ, float dBdyX, float dBdyZ,
                                              float dBdzX, float dBdzY) {
    PmlCoeffs const k = pmlCoeffs(p.tabE, a.g, x, y, z);
    float e0, e1, e2;
    if (k.b[0] * k.b[1] * k.b[2] == 1.0f) {  // outside the layer: kEFull
        e0 = a.E[0][i] + (dBdyZ - dBdzY) * a.c2 * a.dt;
        e1 = a.E[1][i] + (dBdzX - dBdxZ) * a.c2 * a.dt;
        e2 = a.E[2][i] + (dBdxY - dBdyX) * a.c2 * a.dt;
    } else {
        float const yx = k.b[0] * p.psiE[0][i] + k.c[0] * dBdxZ;
        float const zx = k.b[0] * p.psiE[1][i] + k.c[0] * dBdxY;
        float const xy = k.b[1] * p.psiE[2][i] + k.c[1] * dBdyZ;
        float const zy = k.b[1] * p.psiE[3][i] + k.c[1] * dBdyX;
        float const xz = k.b[2] * p.psiE[4][i] + k.c[2] * dBdzY;
        float const yz = k.b[2] * p.psiE[5][i] + k.c[2] * dBdzX;
        p.psiE[0][i] = yx;
        p.psiE[1][i] = zx;
        p.psiE[2][i] = xy;
        p.psiE[3][i] = zy;
        p.psiE[4][i] = xz;
        p.psiE[5][i] = yz;
        float const c2dt = a.c2 * a.dt;
        e0 = a.E[0][i] + c2dt * (divKappa(dBdyZ, k.kappa[1]) - divKappa(dBdzY, k.kappa[2]) + xy - xz);
        e1 = a.E[1][i] + c2dt * (divKappa(dBdzX, k.kappa[2]) - divKappa(dBdxZ, k.kappa[0]) + yz - yx);
        e2 = a.E[2][i] + c2dt * (divKappa(dBdxY, k.kappa[0]) - divKappa(dBdyX, k.kappa[1]) + zx - zy);
    }
    if (ADDJ) {
        e0 = e0 + a.coeff * a.J[0][i];
        e1 = e1 + a.coeff * a.J[1][i];
        e2 = e2 + a.coeff * a.J[2][i];
    }
    Eout[0][i] = e0;
    Eout[1][i] = e1;
    Eout[2][i] = e2;
}

template <int SOLVER>
__global__ __launch_bounds__(256) void kBHalfPml(FieldArgs a, PmlArgs p, int updatePsi) {
    if (halted(a.g)) return;
    size_t const n = (size_t)a.g.nx * a.g.ny * a.g.nz;
    size_t const i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    bHalfPmlAt<SOLVER>(a, p, updatePsi, (int)(i % a.g.nx), (int)((i / a.g.nx) % a.g.ny),
                       (int)(i / ((size_t)a.g.nx * a.g.ny)));
}

// the E full step of every cell with the PML; ADDJ: J -> E in the same pass,
// into Eout. Cells of the inner box skip the coefficient lookup.
template <bool ADDJ>
__device__ __forceinline__ void eFullPmlCell(const FieldArgs& a, const PmlArgs& p, float* const* Eout, size_t i) {
    int const x = (int)(i % a.g.nx), y = (int)((i / a.g.nx) % a.g.ny), z = (int)(i / ((size_t)a.g.nx * a.g.ny));
    Idx I{a.g, x, y, z};
    float const dBdxY = bwd(a.B[1], I, 0, a.cell[0]), dBdxZ = bwd(a.B[2], I, 0, a.cell[0]);
    float const dBdyX = bwd(a.B[0], I, 1, a.cell[1]), dBdyZ = bwd(a.B[2], I, 1, a.cell[1]);
    float const dBdzX = bwd(a.B[0], I, 2, a.cell[2]), dBdzY = bwd(a.B[1], I, 2, a.cell[2]);
    if (!innerCell(a, x, y, z)) {
        eFullPmlApply<ADDJ>(a, p, Eout, x, y, z, i, dBdxY, dBdxZ, dBdyX, dBdyZ, dBdzX, dBdzY);
        return;
    }
    float e0 = a.E[0][i] + (dBdyZ - dBdzY) * a.c2 * a.dt;
    float e1 = a.E[1][i] + (dBdzX - dBdxZ) * a.c2 * a.dt;
    float e2 = a.E[2][i] + (dBdxY - dBdyX) * a.c2 * a.dt;
    if (ADDJ) {
        e0 = e0 + a.coeff * a.J[0][i];
        e1 = e1 + a.coeff * a.J[1][i];
        e2 = e2 + a.coeff * a.J[2][i];
    }
    Eout[0][i] = e0;
    Eout[1][i] = e1;
    Eout[2][i] = e2;
}

__global__ __launch_bounds__(256) void kEFullPml(FieldArgs a, PmlArgs p) {
    if (halted(a.g)) return;
    size_t const n = (size_t)a.g.nx * a.g.ny * a.g.nz;
    size_t const i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    eFullPmlCell<false>(a, p, a.E, i);
}

// kEFull then kAddCurrent (None) per cell in one pass, the new E into Enew
__global__ __launch_bounds__(256) void kEFullJ(FieldArgs a, float* En0, float* En1, float* En2) {
    if (halted(a.g)) return;
    size_t const n = (size_t)a.g.nx * a.g.ny * a.g.nz;
    size_t const i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    Idx I{a.g, (int)(i % a.g.nx), (int)((i / a.g.nx) % a.g.ny), (int)(i / ((size_t)a.g.nx * a.g.ny))};
    float const cx = bwd(a.B[2], I, 1, a.cell[1]) - bwd(a.B[1], I, 2, a.cell[2]);
    float const cy = bwd(a.B[0], I, 2, a.cell[2]) - bwd(a.B[2], I, 0, a.cell[0]);
    float const cz = bwd(a.B[1], I, 0, a.cell[0]) - bwd(a.B[0], I, 1, a.cell[1]);
    float const e0 = a.E[0][i] + cx * a.c2 * a.dt;
    float const e1 = a.E[1][i] + cy * a.c2 * a.dt;
    float const e2 = a.E[2][i] + cz * a.c2 * a.dt;
    En0[i] = e0 + a.coeff * a.J[0][i];
    En1[i] = e1 + a.coeff * a.J[1][i];
    En2[i] = e2 + a.coeff * a.J[2][i];
}

// the PML's E full step (kEFullPml) and J -> E (None) per cell in one pass into Enew
__global__ __launch_bounds__(256) void kEFullPmlJ(FieldArgs a, PmlArgs p, float* En0, float* En1, float* En2) {
    if (halted(a.g)) return;
    size_t const n = (size_t)a.g.nx * a.g.ny * a.g.nz;
    size_t const i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float* const En[3] = {En0, En1, En2};
    eFullPmlCell<true>(a, p, En, i);
}

// ---------------------------------------------------------------------------
// Fused field step (the same per-cell float operations as kBHalf, kEFull,
// kAddCurrent, kBHalf in that order, so the results are bit-identical).
// kFieldBeforeAdd: per 8x8x4 tile, B' = B - curlE(E)*0.5*dt on the tile and
// its lower halo (kept in LDS only), then E'' = (E + curlB(B')*c2*dt) +
// coeff*J, written to Enew. kFieldAfter: B'' = (B - curlE(E)*0.5*dt) -
// curlE(Enew)*0.5*dt per cell, in place. E is read-only in both kernels and
// B is not written by the first, so no block reads a value another block of
// the same launch has overwritten; the caller then swaps E and Enew.
// ---------------------------------------------------------------------------
constexpr int kFTX = 8, kFTY = 8, kFTZ = 4;
constexpr int kFHX = kFTX + 1, kFHY = kFTY + 1, kFHZ = kFTZ + 1;
constexpr int kFHV = kFHX * kFHY * kFHZ;
// E staged for the curl of B' on the tile and its lower halo: 2 cells of halo
// on each side cover the Yee stencil and the Lehe stencil along any axis
// (Lehe/Derivative.hpp:132-145: f(i+2) - f(i-1) along the Cherenkov-free axis,
// shifts of one cell across it)
constexpr int kEH = 2;
constexpr int kEX = kFTX + 2 * kEH, kEY = kFTY + 2 * kEH, kEZ = kFTZ + 2 * kEH;
constexpr int kEV = kEX * kEY * kEZ;
// E and the new E staged for kFieldAfter: halo 1 below, 2 above
constexpr int kAX = kFTX + 3, kAY = kFTY + 3, kAZ = kFTZ + 3;
constexpr int kAV = kAX * kAY * kAZ;

template <int SOLVER>
__device__ __forceinline__ void curlE(const FieldArgs& a, float* const* E, const Idx& I, float& cx, float& cy,
                                      float& cz) {
    cx = derivE<SOLVER>(E[2], I, 1, a) - derivE<SOLVER>(E[1], I, 2, a);
    cy = derivE<SOLVER>(E[0], I, 2, a) - derivE<SOLVER>(E[2], I, 0, a);
    cz = derivE<SOLVER>(E[1], I, 0, a) - derivE<SOLVER>(E[0], I, 1, a);
}

// Derivatives of a component staged in LDS with strides (1, SY, SZ): the
// float operations of fwd/leheDeriv on the global arrays
template <int SY, int SZ>
struct LdsStencil {
    static __device__ __forceinline__ int off(int d) { return d == 0 ? 1 : (d == 1 ? SY : SZ); }
    static __device__ __forceinline__ float fwd(const float* f, int h, int d, float step) {
        return (f[h + off(d)] - f[h]) / step;
    }
    static __device__ __forceinline__ float lehe(const float* f, int h, int d, const FieldArgs& a) {
        int const L = a.lehe_dir;
        float const step = a.cell[d];
        if (d == L) {
            int const o1 = off((d + 1) % 3), o2 = off((d + 2) % 3), o0 = off(d);
            float r = a.lehe_alpha * fwd(f, h, d, step);
            r = r + a.lehe_beta1 * fwd(f, h + o1, d, step);
            r = r + a.lehe_beta1 * fwd(f, h - o1, d, step);
            r = r + a.lehe_beta2 * fwd(f, h + o2, d, step);
            r = r + a.lehe_beta2 * fwd(f, h - o2, d, step);
            r = r + a.lehe_delta * (f[h + 2 * o0] - f[h - o0]) / step;
            return r;
        }
        int const oL = off(L);
        float const beta = 0.125f;
        float const alpha = 1.0f - 2.0f * beta;
        float r = alpha * fwd(f, h, d, step);
        r = r + beta * fwd(f, h + oL, d, step);
        r = r + beta * fwd(f, h - oL, d, step);
        return r;
    }
    template <int SOLVER>
    static __device__ __forceinline__ float deriv(const float* f, int h, int d, const FieldArgs& a) {
        if (SOLVER == PIC_SOLVER_LEHE) return lehe(f, h, d, a);
        return fwd(f, h, d, a.cell[d]);
    }
    // curl of E with components at f0, f1, f2 (each a plane of the staged tile)
    template <int SOLVER>
    static __device__ __forceinline__ void curl(const float* f0, const float* f1, const float* f2, int h,
                                                const FieldArgs& a, float& cx, float& cy, float& cz) {
        cx = deriv<SOLVER>(f2, h, 1, a) - deriv<SOLVER>(f1, h, 2, a);
        cy = deriv<SOLVER>(f0, h, 2, a) - deriv<SOLVER>(f2, h, 0, a);
        cz = deriv<SOLVER>(f1, h, 0, a) - deriv<SOLVER>(f0, h, 1, a);
    }
};

// global index of the cell at tile-relative offset (tx, ty, tz) from (x0, y0, z0), wrapped
__device__ __forceinline__ size_t wrapAt(const GridDev& g, int x0, int y0, int z0, int tx, int ty, int tz) {
    int x = x0 + tx, y = y0 + ty, z = z0 + tz;
    x = x < 0 ? x + g.nx : (x >= g.nx ? x - g.nx : x);
    y = y < 0 ? y + g.ny : (y >= g.ny ? y - g.ny : y);
    z = z < 0 ? z + g.nz : (z >= g.nz ? z - g.nz : z);
    return (size_t)x + (size_t)g.nx * ((size_t)y + (size_t)g.ny * (size_t)z);
}

// beyond an open border, for kernels compiled with OPEN (periodic boxes
// compile the checks away)
template <bool OPEN>
__device__ __forceinline__ bool outsideIf(const GridDev& g, int x, int y, int z) {
    return OPEN && outsideOpen(g, x, y, z);
}

struct FieldLds {
    float E[3][kEV];
    float B[3][kFHV];
};

// tile b of the grid: every thread of the 256-thread workgroup calls it
template <int SOLVER, bool OPEN>
__device__ __forceinline__ void fieldBeforeAddTile(const FieldArgs& a, float* const* Enew, int b, FieldLds& L) {
    int const ntx = a.g.nx / kFTX, nty = a.g.ny / kFTY;
    int const x0 = (b % ntx) * kFTX, y0 = ((b / ntx) % nty) * kFTY, z0 = (b / (ntx * nty)) * kFTZ;
    for (int t = threadIdx.x; t < 3 * kEV; t += blockDim.x) {
        int const c = t / kEV, r = t - c * kEV;
        int const ex = r % kEX, ey = (r / kEX) % kEY, ez = r / (kEX * kEY);
        bool const out = outsideIf<OPEN>(a.g, x0 + ex - kEH, y0 + ey - kEH, z0 + ez - kEH);
        L.E[c][r] = out ? 0.0f : a.E[c][wrapAt(a.g, x0, y0, z0, ex - kEH, ey - kEH, ez - kEH)];
    }
    __syncthreads();
    using SE = LdsStencil<kEX, kEX * kEY>;
    for (int t = threadIdx.x; t < kFHV; t += blockDim.x) {
        int const hx = t % kFHX, hy = (t / kFHX) % kFHY, hz = t / (kFHX * kFHY);
        // B' at tile offset (hx-1, hy-1, hz-1)
        int const he = (hx - 1 + kEH) + kEX * ((hy - 1 + kEH) + kEY * (hz - 1 + kEH));
        size_t const i = wrapAt(a.g, x0, y0, z0, hx - 1, hy - 1, hz - 1);
        if (outsideIf<OPEN>(a.g, x0 + hx - 1, y0 + hy - 1, z0 + hz - 1)) {  // B beyond an open border stays zero
            L.B[0][t] = L.B[1][t] = L.B[2][t] = 0.0f;
            continue;
        }
        float cx, cy, cz;
        SE::template curl<SOLVER>(L.E[0], L.E[1], L.E[2], he, a, cx, cy, cz);
        L.B[0][t] = a.B[0][i] - cx * 0.5f * a.dt;
        L.B[1][t] = a.B[1][i] - cy * 0.5f * a.dt;
        L.B[2][t] = a.B[2][i] - cz * 0.5f * a.dt;
    }
    __syncthreads();
    int const lx = threadIdx.x % kFTX, ly = (threadIdx.x / kFTX) % kFTY, lz = threadIdx.x / (kFTX * kFTY);
    size_t const i = (size_t)(x0 + lx) + (size_t)a.g.nx * ((size_t)(y0 + ly) + (size_t)a.g.ny * (size_t)(z0 + lz));
    int const h = (lx + 1) + kFHX * ((ly + 1) + kFHY * (lz + 1));
    int const he = (lx + kEH) + kEX * ((ly + kEH) + kEY * (lz + kEH));
    constexpr int SY = kFHX, SZ = kFHX * kFHY;
    // backward differences of B' (BackwardDerivative.hpp:44-65), curl as kEFull
    float const cx = (L.B[2][h] - L.B[2][h - SY]) / a.cell[1] - (L.B[1][h] - L.B[1][h - SZ]) / a.cell[2];
    float const cy = (L.B[0][h] - L.B[0][h - SZ]) / a.cell[2] - (L.B[2][h] - L.B[2][h - 1]) / a.cell[0];
    float const cz = (L.B[1][h] - L.B[1][h - 1]) / a.cell[0] - (L.B[0][h] - L.B[0][h - SY]) / a.cell[1];
    float e0 = L.E[0][he] + cx * a.c2 * a.dt;
    float e1 = L.E[1][he] + cy * a.c2 * a.dt;
    float e2 = L.E[2][he] + cz * a.c2 * a.dt;
    e0 = e0 + a.coeff * a.J[0][i];
    e1 = e1 + a.coeff * a.J[1][i];
    e2 = e2 + a.coeff * a.J[2][i];
    Enew[0][i] = e0;
    Enew[1][i] = e1;
    Enew[2][i] = e2;
}

// ---------------------------------------------------------------------------
// Yee: z-marching form. A workgroup owns an xy tile of kZX x kZY cells and a
// chunk of zc planes; it walks up in z with two planes of E and of B' in
// LDS (each with the one-cell xy halo the stencil needs), so E is read once
// per plane plus the halo (two extra planes per chunk) and nothing is
// recomputed along z. The chunk is kZC for ensembles (many members fill the
// chip) and is halved for a single small grid until the launch has
// kZFill workgroups, down to kZCMin planes.
// ---------------------------------------------------------------------------
constexpr int kZX = 32, kZY = 8, kZC = 16, kZCMin = 2, kZFill = 2048;
constexpr int kZEX = kZX + 2, kZEY = kZY + 2;  // E plane: x, y in [-1, T+1)
constexpr int kZBX = kZX + 1, kZBY = kZY + 1;  // B' plane: x, y in [-1, T)

// periodic index for any v (a partial tile near the box edge may reach more than one box length away)
__device__ __forceinline__ int wrap1(int v, int n) {
    v %= n;
    return v < 0 ? v + n : v;
}

struct ZTile {
    int x0, y0, z0, z1;
};

__device__ __forceinline__ ZTile zTile(const GridDev& g, int b, int zc) {
    int const ntx = (g.nx + kZX - 1) / kZX, nty = (g.ny + kZY - 1) / kZY;
    ZTile t;
    t.x0 = (b % ntx) * kZX;
    t.y0 = ((b / ntx) % nty) * kZY;
    int const c = b / (ntx * nty);
    t.z0 = c * zc;
    t.z1 = min(t.z0 + zc, g.nz);
    return t;
}

__host__ __device__ __forceinline__ int zTileCount(const GridDev& g, int zc) {
    return ((g.nx + kZX - 1) / kZX) * ((g.ny + kZY - 1) / kZY) * ((g.nz + zc - 1) / zc);
}

struct YeeBeforeLds {
    float E[2][3][kZEY][kZEX];
    float B[2][3][kZBY][kZBX];
};

// update_beforeCurrent + addCurrent for one z-marching tile (Yee)
template <bool OPEN>
__device__ __forceinline__ void yeeBeforeAddTile(const FieldArgs& a, float* const* Enew, int b, int zc,
                                                 YeeBeforeLds& L) {
    ZTile const T = zTile(a.g, b, zc);
    int const nx = a.g.nx, ny = a.g.ny, nz = a.g.nz;
    size_t const plane = (size_t)nx * ny;
    // per thread, once per tile: the wrapped xy offsets of the plane elements
    // it stages and updates, -1 beyond an open border (the value there is zero)
    constexpr int NE = (3 * kZEX * kZEY + 255) / 256;
    constexpr int NB = (kZBX * kZBY + 255) / 256;
    int eoff[NE];
    for (int j = 0; j < NE; ++j) {
        int const t = threadIdx.x + 256 * j;
        int const r = t % (kZEX * kZEY);
        int const x = T.x0 + r % kZEX - 1, y = T.y0 + r / kZEX - 1;
        eoff[j] = outsideIf<OPEN>(a.g, x, y, 0) ? -1 : wrap1(x, nx) + nx * wrap1(y, ny);
    }
    int boff[NB];
    for (int j = 0; j < NB; ++j) {
        int const t = threadIdx.x + 256 * j;
        int const x = T.x0 + t % kZBX - 1, y = T.y0 + t / kZBX - 1;
        boff[j] = outsideIf<OPEN>(a.g, x, y, 0) ? -1 : wrap1(x, nx) + nx * wrap1(y, ny);
    }
    auto planeOut = [&](int k) { return OPEN && (a.g.open & 4) && (k < 0 || k >= nz); };
    float* const Ebase = &L.E[0][0][0][0];
    auto fetchE = [&](int k, float (&v)[NE]) {  // this thread's elements of E plane k (wrapped)
        size_t const zoff = plane * (size_t)wrap1(k, nz);
        bool const zout = planeOut(k);
#pragma unroll
        for (int j = 0; j < NE; ++j) {
            int const t = threadIdx.x + 256 * j;
            v[j] = 0.0f;
            if (t < 3 * kZEX * kZEY && !(OPEN && (zout || eoff[j] < 0)))
                v[j] = a.E[t / (kZEX * kZEY)][zoff + (size_t)eoff[j]];
        }
    };
    auto storeE = [&](int k, const float (&v)[NE]) {  // into slot k & 1
        float* const dst = Ebase + (k & 1) * (3 * kZEX * kZEY);
#pragma unroll
        for (int j = 0; j < NE; ++j) {
            int const t = threadIdx.x + 256 * j;
            if (t < 3 * kZEX * kZEY) dst[t] = v[j];
        }
    };
    // B' = B - curlE(E)*0.5*dt at plane k on x, y in [-1, T): forward differences
    auto bPrime = [&](int k) {
        size_t const zoff = plane * (size_t)wrap1(k, nz);
        bool const zout = planeOut(k);
        int const s = k & 1, s1 = (k + 1) & 1;
#pragma unroll
        for (int j = 0; j < NB; ++j) {
            int const t = threadIdx.x + 256 * j;
            if (t >= kZBX * kZBY) continue;
            int const bx = t % kZBX, by = t / kZBX;  // cell (bx-1, by-1) of the tile
            if (OPEN && (zout || boff[j] < 0)) {  // B beyond an open border stays zero
                L.B[s][0][by][bx] = L.B[s][1][by][bx] = L.B[s][2][by][bx] = 0.0f;
                continue;
            }
            int const ex = bx, ey = by;              // its E-plane index (E plane starts at -1)
            float const e0 = L.E[s][0][ey][ex], e1 = L.E[s][1][ey][ex], e2 = L.E[s][2][ey][ex];
            float const dEzdy = (L.E[s][2][ey + 1][ex] - e2) / a.cell[1];
            float const dEydz = (L.E[s1][1][ey][ex] - e1) / a.cell[2];
            float const dExdz = (L.E[s1][0][ey][ex] - e0) / a.cell[2];
            float const dEzdx = (L.E[s][2][ey][ex + 1] - e2) / a.cell[0];
            float const dEydx = (L.E[s][1][ey][ex + 1] - e1) / a.cell[0];
            float const dExdy = (L.E[s][0][ey + 1][ex] - e0) / a.cell[1];
            float const cx = dEzdy - dEydz;
            float const cy = dExdz - dEzdx;
            float const cz = dEydx - dExdy;
            size_t const gi = zoff + (size_t)boff[j];
            L.B[s][0][by][bx] = a.B[0][gi] - cx * 0.5f * a.dt;
            L.B[s][1][by][bx] = a.B[1][gi] - cy * 0.5f * a.dt;
            L.B[s][2][by][bx] = a.B[2][gi] - cz * 0.5f * a.dt;
        }
    };
    int const lx = threadIdx.x % kZX, ly = threadIdx.x / kZX;
    bool const inside = T.x0 + lx < nx && T.y0 + ly < ny;
    size_t const ixy = (size_t)(T.x0 + lx) + (size_t)nx * (T.y0 + ly);
    float next[NE];  // plane k+1, fetched while plane k-1 was being computed
    fetchE(T.z0 - 1, next);
    storeE(T.z0 - 1, next);
    fetchE(T.z0, next);
    for (int k = T.z0 - 1; k < T.z1; ++k) {
        __syncthreads();  // slot (k+1)&1 is no longer read
        storeE(k + 1, next);
        if (k + 2 <= T.z1) fetchE(k + 2, next);  // in flight during this step
        __syncthreads();
        bPrime(k);
        __syncthreads();
        if (k >= T.z0 && inside) {
            int const s = k & 1, sm = (k - 1) & 1;
            int const bx = lx + 1, by = ly + 1;
            // backward differences of B' (BackwardDerivative.hpp:44-65), curl as kEFull
            float const cx = (L.B[s][2][by][bx] - L.B[s][2][by - 1][bx]) / a.cell[1] -
                             (L.B[s][1][by][bx] - L.B[sm][1][by][bx]) / a.cell[2];
            float const cy = (L.B[s][0][by][bx] - L.B[sm][0][by][bx]) / a.cell[2] -
                             (L.B[s][2][by][bx] - L.B[s][2][by][bx - 1]) / a.cell[0];
            float const cz = (L.B[s][1][by][bx] - L.B[s][1][by][bx - 1]) / a.cell[0] -
                             (L.B[s][0][by][bx] - L.B[s][0][by - 1][bx]) / a.cell[1];
            size_t const i = ixy + plane * k;
            float e0 = L.E[s][0][ly + 1][lx + 1] + cx * a.c2 * a.dt;
            float e1 = L.E[s][1][ly + 1][lx + 1] + cy * a.c2 * a.dt;
            float e2 = L.E[s][2][ly + 1][lx + 1] + cz * a.c2 * a.dt;
            e0 = e0 + a.coeff * a.J[0][i];
            e1 = e1 + a.coeff * a.J[1][i];
            e2 = e2 + a.coeff * a.J[2][i];
            Enew[0][i] = e0;
            Enew[1][i] = e1;
            Enew[2][i] = e2;
        }
    }
}

struct YeeAfterLds {
    float E[2][3][kZY + 1][kZX + 1];
    float En[2][3][kZY + 1][kZX + 1];
};

// update_afterCurrent for one z-marching tile (Yee): B'' = (B - curlE(E)*0.5*dt) - curlE(Enew)*0.5*dt;
// with TWO = false one B half step alone, B' = B - curlE(E)*0.5*dt (kBHalf's arithmetic)
template <bool OPEN, bool TWO, bool PML = false>
__device__ __forceinline__ void yeeAfterTile(const FieldArgs& a, float* const* Enew, int b, int zc, YeeAfterLds& L,
                                             const PmlArgs* p = nullptr, int updatePsi = 0) {
    ZTile const T = zTile(a.g, b, zc);
    int const nx = a.g.nx, ny = a.g.ny, nz = a.g.nz;
    size_t const plane = (size_t)nx * ny;
    constexpr int PX = kZX + 1, PY = kZY + 1;
    constexpr int NE = (3 * PX * PY + 255) / 256;
    int eoff[NE];
    for (int j = 0; j < NE; ++j) {
        int const t = threadIdx.x + 256 * j;
        int const r = t % (PX * PY);
        int const x = T.x0 + r % PX, y = T.y0 + r / PX;
        eoff[j] = outsideIf<OPEN>(a.g, x, y, 0) ? -1 : wrap1(x, nx) + nx * wrap1(y, ny);
    }
    float* const Eb = &L.E[0][0][0][0];
    float* const Nb = &L.En[0][0][0][0];
    // this thread's elements of planes k of E and Enew on x, y in [0, T+1)
    auto fetch = [&](int k, float (&ve)[NE], float (&vn)[NE]) {
        size_t const zoff = plane * (size_t)wrap1(k, nz);
        bool const zout = OPEN && (a.g.open & 4) && (k < 0 || k >= nz);
#pragma unroll
        for (int j = 0; j < NE; ++j) {
            int const t = threadIdx.x + 256 * j;
            ve[j] = vn[j] = 0.0f;
            if (t < 3 * PX * PY && !(OPEN && (zout || eoff[j] < 0))) {
                int const c = t / (PX * PY);
                size_t const gi = zoff + (size_t)eoff[j];
                ve[j] = a.E[c][gi];
                if (TWO) vn[j] = Enew[c][gi];
            }
        }
    };
    auto store = [&](int k, const float (&ve)[NE], const float (&vn)[NE]) {
        int const so = (k & 1) * (3 * PX * PY);
#pragma unroll
        for (int j = 0; j < NE; ++j) {
            int const t = threadIdx.x + 256 * j;
            if (t < 3 * PX * PY) {
                Eb[so + t] = ve[j];
                if (TWO) Nb[so + t] = vn[j];
            }
        }
    };
    int const lx = threadIdx.x % kZX, ly = threadIdx.x / kZX;
    bool const inside = T.x0 + lx < nx && T.y0 + ly < ny;
    size_t const ixy = (size_t)(T.x0 + lx) + (size_t)nx * (T.y0 + ly);
    float nextE[NE], nextN[NE];  // plane k+1, fetched while plane k-1 was being computed
    fetch(T.z0, nextE, nextN);
    store(T.z0, nextE, nextN);
    fetch(T.z0 + 1, nextE, nextN);
    for (int k = T.z0; k < T.z1; ++k) {
        __syncthreads();
        store(k + 1, nextE, nextN);
        if (k + 2 <= T.z1) fetch(k + 2, nextE, nextN);  // in flight during this step
        __syncthreads();
        if (PML && inside) {
            // B half step with the PML (never with TWO): the six forward
            // differences of the curl for every cell, the update per cell
            int const s = k & 1, s1 = (k + 1) & 1;
            float const f0 = L.E[s][0][ly][lx], f1 = L.E[s][1][ly][lx], f2 = L.E[s][2][ly][lx];
            float const dxY = (L.E[s][1][ly][lx + 1] - f1) / a.cell[0], dxZ = (L.E[s][2][ly][lx + 1] - f2) / a.cell[0];
            float const dyX = (L.E[s][0][ly + 1][lx] - f0) / a.cell[1], dyZ = (L.E[s][2][ly + 1][lx] - f2) / a.cell[1];
            float const dzX = (L.E[s1][0][ly][lx] - f0) / a.cell[2], dzY = (L.E[s1][1][ly][lx] - f1) / a.cell[2];
            size_t const i = ixy + plane * k;
            if (!innerCell(a, T.x0 + lx, T.y0 + ly, k)) {
                bHalfPmlApply(a, *p, updatePsi, T.x0 + lx, T.y0 + ly, k, i, dxY, dxZ, dyX, dyZ, dzX, dzY);
            } else {
                a.B[0][i] = a.B[0][i] - (dyZ - dzY) * 0.5f * a.dt;
                a.B[1][i] = a.B[1][i] - (dzX - dxZ) * 0.5f * a.dt;
                a.B[2][i] = a.B[2][i] - (dxY - dyX) * 0.5f * a.dt;
            }
        } else if (inside) {
            int const s = k & 1, s1 = (k + 1) & 1;
            auto curl = [&](float (*F)[kZY + 1][kZX + 1], float (*G)[kZY + 1][kZX + 1], float& cx, float& cy,
                            float& cz) {
                // F: plane k, G: plane k+1 (forward differences, curl as kBHalf)
                float const f0 = F[0][ly][lx], f1 = F[1][ly][lx], f2 = F[2][ly][lx];
                cx = (F[2][ly + 1][lx] - f2) / a.cell[1] - (G[1][ly][lx] - f1) / a.cell[2];
                cy = (G[0][ly][lx] - f0) / a.cell[2] - (F[2][ly][lx + 1] - f2) / a.cell[0];
                cz = (F[1][ly][lx + 1] - f1) / a.cell[0] - (F[0][ly + 1][lx] - f0) / a.cell[1];
            };
            size_t const i = ixy + plane * k;
            float cx, cy, cz;
            curl(L.E[s], L.E[s1], cx, cy, cz);
            float const b0 = a.B[0][i] - cx * 0.5f * a.dt;
            float const b1 = a.B[1][i] - cy * 0.5f * a.dt;
            float const b2 = a.B[2][i] - cz * 0.5f * a.dt;
            if (!TWO) {
                a.B[0][i] = b0;
                a.B[1][i] = b1;
                a.B[2][i] = b2;
                continue;
            }
            curl(L.En[s], L.En[s1], cx, cy, cz);
            a.B[0][i] = b0 - cx * 0.5f * a.dt;
            a.B[1][i] = b1 - cy * 0.5f * a.dt;
            a.B[2][i] = b2 - cz * 0.5f * a.dt;
        }
    }
}

// ---------------------------------------------------------------------------
// Lehe: z-marching B half step. The Lehe curl of E (Lehe/Derivative.hpp:65-228)
// reaches one cell below and two above along the Cherenkov-free axis and one
// cell across it, so E is staged in a ring of four planes (k-1 .. k+2) with
// an xy halo of two cells. The float operations per cell are those of
// leheDeriv/kBHalf. The Lehe field step writes B' back rather than keeping it
// in LDS as the Yee form does: each Lehe curl costs about 24 correctly
// rounded divisions per cell, and the fused pair would take the first one twice.
// ---------------------------------------------------------------------------
constexpr int kLH = 2;
constexpr int kLEX = kZX + 2 * kLH, kLEY = kZY + 2 * kLH;  // E plane: x, y in [-2, T+2)

// staged field planes: P[k & 3][component][y + kLH][x + kLH]
using LehePlanes = float[4][3][kLEY][kLEX];

__device__ __forceinline__ float lAt(const LehePlanes& P, int c, int x, int y, int k) {
    return P[k & 3][c][y + kLH][x + kLH];
}

// fwd/leheDeriv of component c along d at tile cell (x, y) of plane k
__device__ __forceinline__ float lFwd(const LehePlanes& P, int c, int d, int x, int y, int k, float step) {
    return (lAt(P, c, x + (d == 0), y + (d == 1), k + (d == 2)) - lAt(P, c, x, y, k)) / step;
}

__device__ __forceinline__ float lDeriv(const LehePlanes& P, int c, int d, int x, int y, int k, const FieldArgs& a) {
    int const L = a.lehe_dir;
    float const step = a.cell[d];
    if (d == L) {
        int const d1 = (d + 1) % 3, d2 = (d + 2) % 3;
        int const u1x = d1 == 0, u1y = d1 == 1, u1z = d1 == 2;
        int const u2x = d2 == 0, u2y = d2 == 1, u2z = d2 == 2;
        int const u0x = d == 0, u0y = d == 1, u0z = d == 2;
        float r = a.lehe_alpha * lFwd(P, c, d, x, y, k, step);
        r = r + a.lehe_beta1 * lFwd(P, c, d, x + u1x, y + u1y, k + u1z, step);
        r = r + a.lehe_beta1 * lFwd(P, c, d, x - u1x, y - u1y, k - u1z, step);
        r = r + a.lehe_beta2 * lFwd(P, c, d, x + u2x, y + u2y, k + u2z, step);
        r = r + a.lehe_beta2 * lFwd(P, c, d, x - u2x, y - u2y, k - u2z, step);
        r = r + a.lehe_delta * (lAt(P, c, x + 2 * u0x, y + 2 * u0y, k + 2 * u0z) - lAt(P, c, x - u0x, y - u0y, k - u0z)) /
                    step;
        return r;
    }
    int const ulx = L == 0, uly = L == 1, ulz = L == 2;
    float const beta = 0.125f;
    float const alpha = 1.0f - 2.0f * beta;
    float r = alpha * lFwd(P, c, d, x, y, k, step);
    r = r + beta * lFwd(P, c, d, x + ulx, y + uly, k + ulz, step);
    r = r + beta * lFwd(P, c, d, x - ulx, y - uly, k - ulz, step);
    return r;
}

__device__ __forceinline__ void lCurl(const LehePlanes& P, int x, int y, int k, const FieldArgs& a, float& cx, float& cy,
                                      float& cz) {
    cx = lDeriv(P, 2, 1, x, y, k, a) - lDeriv(P, 1, 2, x, y, k, a);
    cy = lDeriv(P, 0, 2, x, y, k, a) - lDeriv(P, 2, 0, x, y, k, a);
    cz = lDeriv(P, 1, 0, x, y, k, a) - lDeriv(P, 0, 1, x, y, k, a);
}

// per-thread wrapped xy offsets of the staged plane elements (-1: beyond an open border)
template <bool OPEN, int N>
__device__ __forceinline__ void lehePlaneOffsets(const GridDev& g, const ZTile& T, int (&off)[N]) {
    for (int j = 0; j < N; ++j) {
        int const t = threadIdx.x + 256 * j;
        int const r = t % (kLEX * kLEY);
        int const x = T.x0 + r % kLEX - kLH, y = T.y0 + r / kLEX - kLH;
        off[j] = outsideIf<OPEN>(g, x, y, 0) ? -1 : wrap1(x, g.nx) + g.nx * wrap1(y, g.ny);
    }
}

// this thread's elements of plane k of the three components of F (zero beyond an open border)
template <bool OPEN, int N>
__device__ __forceinline__ void lehePlaneFetch(const GridDev& g, float* const* F, const int (&off)[N], int k,
                                               float (&v)[N]) {
    size_t const zoff = (size_t)g.nx * g.ny * (size_t)wrap1(k, g.nz);
    bool const zout = OPEN && (g.open & 4) && (k < 0 || k >= g.nz);
#pragma unroll
    for (int j = 0; j < N; ++j) {
        int const t = threadIdx.x + 256 * j;
        v[j] = 0.0f;
        if (t < 3 * kLEX * kLEY && !(OPEN && (zout || off[j] < 0)))
            v[j] = F[t / (kLEX * kLEY)][zoff + (size_t)off[j]];
    }
}

// fetched elements into the ring slot of plane k
template <int N>
__device__ __forceinline__ void lehePlaneStore(const float (&v)[N], int k, LehePlanes& P) {
    float* const dst = &P[k & 3][0][0][0];
#pragma unroll
    for (int j = 0; j < N; ++j) {
        int const t = threadIdx.x + 256 * j;
        if (t < 3 * kLEX * kLEY) dst[t] = v[j];
    }
}

template <bool OPEN, int N>
__device__ __forceinline__ void lehePlaneLoad(const GridDev& g, float* const* F, const int (&off)[N], int k,
                                              LehePlanes& P) {
    float v[N];
    lehePlaneFetch<OPEN>(g, F, off, k, v);
    lehePlaneStore(v, k, P);
}

// one B half step for one z-marching tile (Lehe): B' = B - curlE(E)*0.5*dt,
// the arithmetic of kBHalf
template <bool OPEN, bool PML>
__device__ __forceinline__ void leheBHalfTile(const FieldArgs& a, const PmlArgs& p, int updatePsi, int b, int zc,
                                              LehePlanes& LE) {
    ZTile const T = zTile(a.g, b, zc);
    int const nx = a.g.nx, ny = a.g.ny;
    size_t const plane = (size_t)nx * ny;
    constexpr int NE = (3 * kLEX * kLEY + 255) / 256;
    int eoff[NE];
    lehePlaneOffsets<OPEN>(a.g, T, eoff);
    for (int k = T.z0 - 1; k <= T.z0 + 1; ++k) lehePlaneLoad<OPEN>(a.g, a.E, eoff, k, LE);
    int const lx = threadIdx.x % kZX, ly = threadIdx.x / kZX;
    bool const inside = T.x0 + lx < nx && T.y0 + ly < ny;
    size_t const ixy = (size_t)(T.x0 + lx) + (size_t)nx * (T.y0 + ly);
    float nextE[NE];  // plane k+2, fetched while plane k-1 was being computed
    lehePlaneFetch<OPEN>(a.g, a.E, eoff, T.z0 + 2, nextE);
    for (int k = T.z0; k < T.z1; ++k) {
        __syncthreads();  // the slot of plane k+2 (plane k-2) is no longer read
        lehePlaneStore(nextE, k + 2, LE);
        if (k + 3 <= T.z1 + 1) lehePlaneFetch<OPEN>(a.g, a.E, eoff, k + 3, nextE);  // in flight during this step
        __syncthreads();
        if (PML && inside) {
            // the six derivatives of the curl for every cell, so that layer
            // and inner cells of a wave share them; only the update branches
            float const dxY = lDeriv(LE, 1, 0, lx, ly, k, a), dxZ = lDeriv(LE, 2, 0, lx, ly, k, a);
            float const dyX = lDeriv(LE, 0, 1, lx, ly, k, a), dyZ = lDeriv(LE, 2, 1, lx, ly, k, a);
            float const dzX = lDeriv(LE, 0, 2, lx, ly, k, a), dzY = lDeriv(LE, 1, 2, lx, ly, k, a);
            size_t const i = ixy + plane * k;
            if (!innerCell(a, T.x0 + lx, T.y0 + ly, k)) {
                bHalfPmlApply(a, p, updatePsi, T.x0 + lx, T.y0 + ly, k, i, dxY, dxZ, dyX, dyZ, dzX, dzY);
            } else {
                a.B[0][i] = a.B[0][i] - (dyZ - dzY) * 0.5f * a.dt;
                a.B[1][i] = a.B[1][i] - (dzX - dxZ) * 0.5f * a.dt;
                a.B[2][i] = a.B[2][i] - (dxY - dyX) * 0.5f * a.dt;
            }
        } else if (inside) {
            size_t const i = ixy + plane * k;
            float cx, cy, cz;
            lCurl(LE, lx, ly, k, a, cx, cy, cz);
            a.B[0][i] = a.B[0][i] - cx * 0.5f * a.dt;
            a.B[1][i] = a.B[1][i] - cy * 0.5f * a.dt;
            a.B[2][i] = a.B[2][i] - cz * 0.5f * a.dt;
        }
    }
}

struct AfterLds {
    float E[3][kAV];
    float En[3][kAV];
};

// B'' = (B - curlE(E)*0.5*dt) - curlE(Enew)*0.5*dt on tile b
template <int SOLVER, bool OPEN>
__device__ __forceinline__ void fieldAfterTile(const FieldArgs& a, float* const* Enew, int b, AfterLds& L) {
    int const ntx = a.g.nx / kFTX, nty = a.g.ny / kFTY;
    int const x0 = (b % ntx) * kFTX, y0 = ((b / ntx) % nty) * kFTY, z0 = (b / (ntx * nty)) * kFTZ;
    for (int t = threadIdx.x; t < 3 * kAV; t += blockDim.x) {
        int const c = t / kAV, r = t - c * kAV;
        int const ex = r % kAX, ey = (r / kAX) % kAY, ez = r / (kAX * kAY);
        size_t const gi = wrapAt(a.g, x0, y0, z0, ex - 1, ey - 1, ez - 1);
        bool const out = outsideIf<OPEN>(a.g, x0 + ex - 1, y0 + ey - 1, z0 + ez - 1);
        L.E[c][r] = out ? 0.0f : a.E[c][gi];
        L.En[c][r] = out ? 0.0f : Enew[c][gi];
    }
    __syncthreads();
    using SA = LdsStencil<kAX, kAX * kAY>;
    int const lx = threadIdx.x % kFTX, ly = (threadIdx.x / kFTX) % kFTY, lz = threadIdx.x / (kFTX * kFTY);
    size_t const i = (size_t)(x0 + lx) + (size_t)a.g.nx * ((size_t)(y0 + ly) + (size_t)a.g.ny * (size_t)(z0 + lz));
    int const h = (lx + 1) + kAX * ((ly + 1) + kAY * (lz + 1));
    float cx, cy, cz;
    SA::template curl<SOLVER>(L.E[0], L.E[1], L.E[2], h, a, cx, cy, cz);
    float const b0 = a.B[0][i] - cx * 0.5f * a.dt;
    float const b1 = a.B[1][i] - cy * 0.5f * a.dt;
    float const b2 = a.B[2][i] - cz * 0.5f * a.dt;
    SA::template curl<SOLVER>(L.En[0], L.En[1], L.En[2], h, a, cx, cy, cz);
    a.B[0][i] = b0 - cx * 0.5f * a.dt;
    a.B[1][i] = b1 - cy * 0.5f * a.dt;
    a.B[2][i] = b2 - cz * 0.5f * a.dt;
}

template <int SOLVER, bool OPEN>
__global__ __launch_bounds__(256) void kFieldBeforeAdd(FieldArgs a, float* Enew0, float* Enew1, float* Enew2) {
    if (halted(a.g)) return;
    __shared__ FieldLds L;
    float* const En[3] = {Enew0, Enew1, Enew2};
    fieldBeforeAddTile<SOLVER, OPEN>(a, En, blockIdx.x, L);
}

template <bool OPEN>
__global__ __launch_bounds__(256) void kYeeBeforeAdd(FieldArgs a, float* Enew0, float* Enew1, float* Enew2, int zc) {
    if (halted(a.g)) return;
    __shared__ YeeBeforeLds L;
    float* const En[3] = {Enew0, Enew1, Enew2};
    yeeBeforeAddTile<OPEN>(a, En, blockIdx.x, zc, L);
}

template <bool OPEN>
__global__ __launch_bounds__(256) void kYeeAfter(FieldArgs a, float* Enew0, float* Enew1, float* Enew2, int zc) {
    if (halted(a.g)) return;
    __shared__ YeeAfterLds L;
    float* const En[3] = {Enew0, Enew1, Enew2};
    yeeAfterTile<OPEN, true>(a, En, blockIdx.x, zc, L);
}

// one B half step (the stage call kBHalf) in the z-marching form
template <bool OPEN>
__global__ __launch_bounds__(256) void kLeheBHalfZ(FieldArgs a, int zc) {
    if (halted(a.g)) return;
    __shared__ LehePlanes L;
    leheBHalfTile<OPEN, false>(a, PmlArgs{}, 0, blockIdx.x, zc, L);
}

// with the PML: the layer cells (outside the inner box) as kBHalfPml; the
// tiles listed after the plain ones when there is a list
template <bool OPEN>
__global__ __launch_bounds__(256) void kLeheBHalfPmlZ(FieldArgs a, PmlArgs p, int updatePsi, int zc) {
    if (halted(a.g)) return;
    __shared__ LehePlanes L;
    int const b = p.tiles ? p.tiles[p.nPlain + blockIdx.x] : (int)blockIdx.x;
    leheBHalfTile<OPEN, true>(a, p, updatePsi, b, zc, L);
}

// the plain B half step on the listed tiles (no cell of them in a layer)
template <bool OPEN>
__global__ __launch_bounds__(256) void kLeheBHalfListZ(FieldArgs a, const int* __restrict__ tiles, int zc) {
    if (halted(a.g)) return;
    __shared__ LehePlanes L;
    leheBHalfTile<OPEN, false>(a, PmlArgs{}, 0, tiles[blockIdx.x], zc, L);
}

template <bool OPEN>
__global__ __launch_bounds__(256) void kYeeBHalfZ(FieldArgs a, int zc) {
    if (halted(a.g)) return;
    __shared__ YeeAfterLds L;
    yeeAfterTile<OPEN, false>(a, nullptr, blockIdx.x, zc, L);
}

template <bool OPEN>
__global__ __launch_bounds__(256) void kYeeBHalfPmlZ(FieldArgs a, PmlArgs p, int updatePsi, int zc) {
    if (halted(a.g)) return;
    __shared__ YeeAfterLds L;
    int const b = p.tiles ? p.tiles[p.nPlain + blockIdx.x] : (int)blockIdx.x;
    yeeAfterTile<OPEN, false, true>(a, nullptr, b, zc, L, &p, updatePsi);
}

template <bool OPEN>
__global__ __launch_bounds__(256) void kYeeBHalfListZ(FieldArgs a, const int* __restrict__ tiles, int zc) {
    if (halted(a.g)) return;
    __shared__ YeeAfterLds L;
    yeeAfterTile<OPEN, false>(a, nullptr, tiles[blockIdx.x], zc, L);
}

__global__ __launch_bounds__(256) void kYeeBeforeAddBatch(const FieldBatchEntry* __restrict__ table,
                                                          const int2* __restrict__ tiles) {
    __shared__ YeeBeforeLds L;
    int2 const it = tiles[blockIdx.x];
    const FieldBatchEntry& e = table[it.x];
    if (halted(e.a.g)) return;
    yeeBeforeAddTile<false>(e.a, e.Enew, it.y, kZC, L);
}

__global__ __launch_bounds__(256) void kYeeAfterBatch(const FieldBatchEntry* __restrict__ table,
                                                      const int2* __restrict__ tiles) {
    __shared__ YeeAfterLds L;
    int2 const it = tiles[blockIdx.x];
    const FieldBatchEntry& e = table[it.x];
    if (halted(e.a.g)) return;
    yeeAfterTile<false, true>(e.a, e.Enew, it.y, kZC, L);
}

template <int SOLVER, bool OPEN>
__global__ __launch_bounds__(256) void kFieldAfter(FieldArgs a, float* Enew0, float* Enew1, float* Enew2) {
    if (halted(a.g)) return;
    __shared__ AfterLds L;
    float* const En[3] = {Enew0, Enew1, Enew2};
    fieldAfterTile<SOLVER, OPEN>(a, En, blockIdx.x, L);
}

// batched forms for ensembles: items (member, tile)
template <int SOLVER>
__global__ __launch_bounds__(256) void kFieldBeforeAddBatch(const FieldBatchEntry* __restrict__ table,
                                                            const int2* __restrict__ tiles) {
    __shared__ FieldLds L;
    int2 const it = tiles[blockIdx.x];
    const FieldBatchEntry& e = table[it.x];
    if (halted(e.a.g)) return;
    fieldBeforeAddTile<SOLVER, false>(e.a, e.Enew, it.y, L);
}

template <int SOLVER>
__global__ __launch_bounds__(256) void kFieldAfterBatch(const FieldBatchEntry* __restrict__ table,
                                                        const int2* __restrict__ tiles) {
    __shared__ AfterLds L;
    int2 const it = tiles[blockIdx.x];
    const FieldBatchEntry& e = table[it.x];
    if (halted(e.a.g)) return;
    fieldAfterTile<SOLVER, false>(e.a, e.Enew, it.y, L);
}

__global__ __launch_bounds__(256) void kCurrentResetBatch(const FieldBatchEntry* __restrict__ table,
                                                          const int2* __restrict__ blocks) {
    int2 const it = blocks[blockIdx.x];
    const FieldBatchEntry& e = table[it.x];
    if (halted(e.a.g)) return;
    size_t const n = (size_t)e.a.g.nx * e.a.g.ny * e.a.g.nz;
    size_t const i = (size_t)it.y * blockDim.x + threadIdx.x;
    if (i >= n) return;
    e.a.J[0][i] = 0.0f;
    e.a.J[1][i] = 0.0f;
    e.a.J[2][i] = 0.0f;
}

__global__ __launch_bounds__(256) void kAddCurrent(FieldArgs a) {
    if (halted(a.g)) return;
    size_t const n = (size_t)a.g.nx * a.g.ny * a.g.nz;
    size_t const i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    a.E[0][i] = a.E[0][i] + a.coeff * a.J[0][i];
    a.E[1][i] = a.E[1][i] + a.coeff * a.J[1][i];
    a.E[2][i] = a.E[2][i] + a.coeff * a.J[2][i];
}

// Binomial current filter (currentInterpolation/Binomial.hpp:40-104) of one
// component at cell I on a periodic box: corner, edge and face sums in the
// reference's order with weights 1, 2, 4, the centre with 8, times 1/64
__device__ __forceinline__ float binomialJ(const float* J, const Idx& I) {
    auto j = [&](int dx, int dy, int dz) { return J[I.at(dx, dy, dz)]; };
    float const corner = j(-1, -1, -1) + j(+1, -1, -1) + j(-1, +1, -1) + j(+1, +1, -1) + j(-1, -1, +1) +
                         j(+1, -1, +1) + j(-1, +1, +1) + j(+1, +1, +1);
    float const edge = j(-1, -1, 0) + j(+1, -1, 0) + j(-1, +1, 0) + j(+1, +1, 0) + j(-1, 0, -1) + j(+1, 0, -1) +
                       j(-1, 0, +1) + j(+1, 0, +1) + j(0, -1, -1) + j(0, +1, -1) + j(0, -1, +1) + j(0, +1, +1);
    float const face = j(-1, 0, 0) + j(+1, 0, 0) + j(0, -1, 0) + j(0, +1, 0) + j(0, 0, -1) + j(0, 0, +1);
    float const avg = 1.0f * corner + 2.0f * edge + 4.0f * face + 8.0f * j(0, 0, 0);
    return avg * (1.0f / 64.0f);
}

__global__ __launch_bounds__(256) void kAddCurrentBinomial(FieldArgs a) {
    if (halted(a.g)) return;
    size_t const n = (size_t)a.g.nx * a.g.ny * a.g.nz;
    size_t const i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    Idx const I{a.g, (int)(i % a.g.nx), (int)((i / a.g.nx) % a.g.ny), (int)(i / ((size_t)a.g.nx * a.g.ny))};
    a.E[0][i] = a.E[0][i] + a.coeff * binomialJ(a.J[0], I);
    a.E[1][i] = a.E[1][i] + a.coeff * binomialJ(a.J[1], I);
    a.E[2][i] = a.E[2][i] + a.coeff * binomialJ(a.J[2], I);
}

// the filtered current into out (the fused field step then adds it as J)
__global__ __launch_bounds__(256) void kBinomialJ(FieldArgs a, float* o0, float* o1, float* o2) {
    if (halted(a.g)) return;
    size_t const n = (size_t)a.g.nx * a.g.ny * a.g.nz;
    size_t const i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    Idx const I{a.g, (int)(i % a.g.nx), (int)((i / a.g.nx) % a.g.ny), (int)(i / ((size_t)a.g.nx * a.g.ny))};
    o0[i] = binomialJ(a.J[0], I);
    o1[i] = binomialJ(a.J[1], I);
    o2[i] = binomialJ(a.J[2], I);
}

__global__ __launch_bounds__(256) void kFill(float* p, size_t n, float v) {
    size_t const i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

// CurrentReset (simulation/stage/CurrentReset.hpp:39-53) at the start of a step
__global__ __launch_bounds__(256) void kCurrentReset(GridDev g, float* J0, float* J1, float* J2, size_t n) {
    if (halted(g)) return;
    size_t const i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    J0[i] = 0.0f;
    J1[i] = 0.0f;
    J2[i] = 0.0f;
}

// end of a batched step, per member: a halted member counts the step as
// skipped; otherwise an overflow in this step halts the following ones
__global__ __launch_bounds__(64) void kEndStepBatch(const FieldBatchEntry* __restrict__ table, int n) {
    int const i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    StepFlags* f = table[i].a.g.flags;
    if (!f) return;
    if (f->halt)
        f->skipped += 1u;
    else if (f->overflow)
        f->halt = 1u;
}

// EnergyFields::reduceField (plugins/EnergyFields.x.cpp:240-247): per
// component the square in FP32 (squareComponentWise), cast to FP64
// (cast64Bit), summed in FP64. Per block the partial sums of the six
// components (B x, y, z, then E x, y, z) at partials[6 * block + k].
__global__ __launch_bounds__(256) void kEnergy(FieldArgs a, double* partials) {
    __shared__ double sAcc[6][256];
    size_t const n = (size_t)a.g.nx * a.g.ny * a.g.nz;
    double acc[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            float const bv = a.B[c][i], ev = a.E[c][i];
            float const b2 = bv * bv, e2 = ev * ev;
            acc[c] += (double)b2;
            acc[3 + c] += (double)e2;
        }
    }
#pragma unroll
    for (int k = 0; k < 6; ++k) sAcc[k][threadIdx.x] = acc[k];
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) {
#pragma unroll
            for (int k = 0; k < 6; ++k) sAcc[k][threadIdx.x] += sAcc[k][threadIdx.x + s];
        }
        __syncthreads();
    }
    if (threadIdx.x < 6) partials[6 * blockIdx.x + threadIdx.x] = sAcc[threadIdx.x][0];
}

// the six sums over the blocks' partials, block by block in ascending order
__global__ __launch_bounds__(64) void kEnergyFinal(const double* __restrict__ partials, int nblocks, double* out) {
    int const k = threadIdx.x;
    if (k >= 6) return;
    double acc = 0.0;
    for (int b = 0; b < nblocks; ++b) acc += partials[6 * b + k];
    out[k] = acc;
}

__global__ void kCountSum(const uint32_t* c, int n, unsigned long long* out) {
    __shared__ unsigned long long s[256];
    unsigned long long acc = 0;
    for (int i = threadIdx.x; i < n; i += blockDim.x) acc += c[i];
    s[threadIdx.x] = acc;
    __syncthreads();
    for (int k = 128; k > 0; k >>= 1) {
        if ((int)threadIdx.x < k) s[threadIdx.x] += s[threadIdx.x + k];
        __syncthreads();
    }
    if (threadIdx.x == 0) *out = s[0];
}

static inline int blocksFor(size_t n) { return (int)((n + 255) / 256); }

int yeeChunk(const GridDev& g);

hipError_t launchBHalf(int solver, const FieldArgs& a, hipStream_t s) {
    size_t const n = (size_t)a.g.nx * a.g.ny * a.g.nz;
    int const zc = yeeChunk(a.g);  // the z-marching form when the grid fills the chip
    if (zc > 0) {
        unsigned const zt = (unsigned)zTileCount(a.g, zc);
        if (solver == PIC_SOLVER_LEHE) {
            if (a.g.open)
                hipLaunchKernelGGL(kLeheBHalfZ<true>, dim3(zt), dim3(256), 0, s, a, zc);
            else
                hipLaunchKernelGGL(kLeheBHalfZ<false>, dim3(zt), dim3(256), 0, s, a, zc);
        } else {
            if (a.g.open)
                hipLaunchKernelGGL(kYeeBHalfZ<true>, dim3(zt), dim3(256), 0, s, a, zc);
            else
                hipLaunchKernelGGL(kYeeBHalfZ<false>, dim3(zt), dim3(256), 0, s, a, zc);
        }
        return hipGetLastError();
    }
    if (solver == PIC_SOLVER_LEHE)
        hipLaunchKernelGGL(kBHalf<PIC_SOLVER_LEHE>, dim3(blocksFor(n)), dim3(256), 0, s, a);
    else
        hipLaunchKernelGGL(kBHalf<PIC_SOLVER_YEE>, dim3(blocksFor(n)), dim3(256), 0, s, a);
    return hipGetLastError();
}

hipError_t launchBHalfPml(int solver, const FieldArgs& a, const PmlArgs& p, bool updatePsi, hipStream_t s) {
    size_t const n = (size_t)a.g.nx * a.g.ny * a.g.nz;
    int const zc = yeeChunk(a.g);  // the z-marching form when the grid fills the chip
    if (zc > 0) {
        unsigned const zt = p.tiles ? (unsigned)p.nLayer : (unsigned)zTileCount(a.g, zc);
        int const u = updatePsi ? 1 : 0;
        bool const lehe = solver == PIC_SOLVER_LEHE;
        if (p.tiles && p.nPlain > 0) {  // the tiles without a layer cell: the plain kernel
            dim3 const grid((unsigned)p.nPlain);
            if (lehe && a.g.open)
                hipLaunchKernelGGL(kLeheBHalfListZ<true>, grid, dim3(256), 0, s, a, p.tiles, zc);
            else if (lehe)
                hipLaunchKernelGGL(kLeheBHalfListZ<false>, grid, dim3(256), 0, s, a, p.tiles, zc);
            else if (a.g.open)
                hipLaunchKernelGGL(kYeeBHalfListZ<true>, grid, dim3(256), 0, s, a, p.tiles, zc);
            else
                hipLaunchKernelGGL(kYeeBHalfListZ<false>, grid, dim3(256), 0, s, a, p.tiles, zc);
        }
        if (zt > 0) {
            if (lehe && a.g.open)
                hipLaunchKernelGGL(kLeheBHalfPmlZ<true>, dim3(zt), dim3(256), 0, s, a, p, u, zc);
            else if (lehe)
                hipLaunchKernelGGL(kLeheBHalfPmlZ<false>, dim3(zt), dim3(256), 0, s, a, p, u, zc);
            else if (a.g.open)
                hipLaunchKernelGGL(kYeeBHalfPmlZ<true>, dim3(zt), dim3(256), 0, s, a, p, u, zc);
            else
                hipLaunchKernelGGL(kYeeBHalfPmlZ<false>, dim3(zt), dim3(256), 0, s, a, p, u, zc);
        }
        return hipGetLastError();
    }
    if (solver == PIC_SOLVER_LEHE)
        hipLaunchKernelGGL(kBHalfPml<PIC_SOLVER_LEHE>, dim3(blocksFor(n)), dim3(256), 0, s, a, p, updatePsi ? 1 : 0);
    else
        hipLaunchKernelGGL(kBHalfPml<PIC_SOLVER_YEE>, dim3(blocksFor(n)), dim3(256), 0, s, a, p, updatePsi ? 1 : 0);
    return hipGetLastError();
}

hipError_t launchEFullPml(const FieldArgs& a, const PmlArgs& p, hipStream_t s) {
    size_t const n = (size_t)a.g.nx * a.g.ny * a.g.nz;
    hipLaunchKernelGGL(kEFullPml, dim3(blocksFor(n)), dim3(256), 0, s, a, p);
    return hipGetLastError();
}

hipError_t launchEFullPmlJ(const FieldArgs& a, const PmlArgs& p, float* const Enew[3], hipStream_t s) {
    size_t const n = (size_t)a.g.nx * a.g.ny * a.g.nz;
    hipLaunchKernelGGL(kEFullPmlJ, dim3(blocksFor(n)), dim3(256), 0, s, a, p, Enew[0], Enew[1], Enew[2]);
    return hipGetLastError();
}

hipError_t launchEFullJ(const FieldArgs& a, float* const Enew[3], hipStream_t s) {
    size_t const n = (size_t)a.g.nx * a.g.ny * a.g.nz;
    hipLaunchKernelGGL(kEFullJ, dim3(blocksFor(n)), dim3(256), 0, s, a, Enew[0], Enew[1], Enew[2]);
    return hipGetLastError();
}

hipError_t launchEFull(const FieldArgs& a, hipStream_t s) {
    size_t const n = (size_t)a.g.nx * a.g.ny * a.g.nz;
    hipLaunchKernelGGL(kEFull, dim3(blocksFor(n)), dim3(256), 0, s, a);
    return hipGetLastError();
}

// Exponential absorber, one layer (absorber/exponential/Exponential.kernel:63-112):
// the cell at distance d from the border of `axis` (minSide: the low border)
// is multiplied by table[depth - d] for d < depth, each component. Threads
// walk the slab x fastest, so rows along x stay coalesced.
__global__ __launch_bounds__(256) void kAbsorb(float* f0, float* f1, float* f2, GridDev g, int axis, int minSide,
                                               int depth, const float* __restrict__ table) {
    if (halted(g)) return;
    int const sx = axis == 0 ? depth : g.nx, sy = axis == 1 ? depth : g.ny, sz = axis == 2 ? depth : g.nz;
    size_t const t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (size_t)sx * sy * sz) return;
    int c[3] = {(int)(t % sx), (int)((t / sx) % sy), (int)(t / ((size_t)sx * sy))};
    int const n[3] = {g.nx, g.ny, g.nz};
    int const d = c[axis];
    if (!minSide) c[axis] = n[axis] - 1 - d;
    float const a = table[depth - d];
    size_t const i = (size_t)c[0] + (size_t)g.nx * ((size_t)c[1] + (size_t)g.ny * c[2]);
    f0[i] = f0[i] * a;
    f1[i] = f1[i] * a;
    f2[i] = f2[i] * a;
}

hipError_t launchAbsorb(const GridDev& g, float* const F[3], const AbsorbPlan& p, hipStream_t s) {
    for (int l = 0; l < p.n; ++l) {
        const AbsorbLayer& L = p.layer[l];
        size_t const other = L.axis == 0 ? (size_t)g.ny * g.nz : (L.axis == 1 ? (size_t)g.nx * g.nz : (size_t)g.nx * g.ny);
        size_t const cells = other * (size_t)L.depth;
        if (cells == 0) continue;
        hipLaunchKernelGGL(kAbsorb, dim3((unsigned)((cells + 255) / 256)), dim3(256), 0, s, F[0], F[1], F[2], g, L.axis,
                           L.minSide, L.depth, L.table);
    }
    return hipGetLastError();
}

int pmlTileCount(const GridDev& g) {
    int const zc = yeeChunk(g);
    return zc > 0 ? zTileCount(g, zc) : 0;
}

void pmlTileSplit(const GridDev& g, const int in0[3], const int in1[3], int* out, int* nPlain) {
    int const zc = yeeChunk(g);
    int const n = zTileCount(g, zc);
    int const ntx = (g.nx + kZX - 1) / kZX, nty = (g.ny + kZY - 1) / kZY;
    int plain = 0;
    for (int b = 0; b < n; ++b) {
        int const x0 = (b % ntx) * kZX, y0 = ((b / ntx) % nty) * kZY, z0 = (b / (ntx * nty)) * zc;
        int const x1 = min(x0 + kZX, g.nx), y1 = min(y0 + kZY, g.ny), z1 = min(z0 + zc, g.nz);
        bool const inside = x0 >= in0[0] && x1 <= in1[0] && y0 >= in0[1] && y1 <= in1[1] && z0 >= in0[2] &&
                            z1 <= in1[2];
        if (inside) out[plain++] = b;
    }
    *nPlain = plain;
    int k = plain;
    for (int b = 0; b < n; ++b) {
        int const x0 = (b % ntx) * kZX, y0 = ((b / ntx) % nty) * kZY, z0 = (b / (ntx * nty)) * zc;
        int const x1 = min(x0 + kZX, g.nx), y1 = min(y0 + kZY, g.ny), z1 = min(z0 + zc, g.nz);
        bool const inside = x0 >= in0[0] && x1 <= in1[0] && y0 >= in0[1] && y1 <= in1[1] && z0 >= in0[2] &&
                            z1 <= in1[2];
        if (!inside) out[k++] = b;
    }
}

int yeeChunk(const GridDev& g) {
    int zc = kZC;
    while (zc > kZCMin && zTileCount(g, zc) < kZFill) zc /= 2;
    return zTileCount(g, zc) >= kZFill / 2 ? zc : 0;
}

template <bool OPEN>
static hipError_t fieldStepT(int solver, const FieldArgs& a, float* const Enew[3], const AbsorbPlan* ab, hipStream_t s) {
    // update_afterCurrent damps E before its B half step and B after it
    // (FDTDBase.hpp:151-183): between the two kernels and after the second
    bool const absorb = ab && ab->n > 0;
    int const tiles = (a.g.nx / kFTX) * (a.g.ny / kFTY) * (a.g.nz / kFTZ);
    int const zc = yeeChunk(a.g);  // both solvers z-march when the grid fills the chip
    int const zt = zc > 0 ? zTileCount(a.g, zc) : 0;
    if (solver == PIC_SOLVER_LEHE && zc > 0) {
        // Lehe: B' is written back so that each of the expensive Lehe curls is
        // taken once (the fused pair would take the first one twice): B half,
        // E full step and J into Enew, absorber, B half from Enew, absorber
        hipLaunchKernelGGL(kLeheBHalfZ<OPEN>, dim3(zt), dim3(256), 0, s, a, zc);
        size_t const n = (size_t)a.g.nx * a.g.ny * a.g.nz;
        hipLaunchKernelGGL(kEFullJ, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, a, Enew[0], Enew[1], Enew[2]);
        if (absorb) {
            hipError_t const e = launchAbsorb(a.g, Enew, *ab, s);
            if (e != hipSuccess) return e;
        }
        FieldArgs an = a;
        for (int d = 0; d < 3; ++d) an.E[d] = Enew[d];
        hipLaunchKernelGGL(kLeheBHalfZ<OPEN>, dim3(zt), dim3(256), 0, s, an, zc);
        if (absorb) return launchAbsorb(a.g, a.B, *ab, s);
        return hipGetLastError();
    }
    if (solver == PIC_SOLVER_LEHE)
        hipLaunchKernelGGL((kFieldBeforeAdd<PIC_SOLVER_LEHE, OPEN>), dim3(tiles), dim3(256), 0, s, a, Enew[0], Enew[1],
                           Enew[2]);
    else if (zc == 0)  // grids too small for the z-marching form to fill the chip: 3D tiles of 8x8x4
        hipLaunchKernelGGL((kFieldBeforeAdd<PIC_SOLVER_YEE, OPEN>), dim3(tiles), dim3(256), 0, s, a, Enew[0], Enew[1],
                           Enew[2]);
    else
        hipLaunchKernelGGL(kYeeBeforeAdd<OPEN>, dim3(zt), dim3(256), 0, s, a, Enew[0], Enew[1], Enew[2], zc);
    if (absorb) {
        hipError_t const e = launchAbsorb(a.g, Enew, *ab, s);
        if (e != hipSuccess) return e;
    }
    if (solver == PIC_SOLVER_LEHE)
        hipLaunchKernelGGL((kFieldAfter<PIC_SOLVER_LEHE, OPEN>), dim3(tiles), dim3(256), 0, s, a, Enew[0], Enew[1],
                           Enew[2]);
    else if (zc == 0)
        hipLaunchKernelGGL((kFieldAfter<PIC_SOLVER_YEE, OPEN>), dim3(tiles), dim3(256), 0, s, a, Enew[0], Enew[1],
                           Enew[2]);
    else
        hipLaunchKernelGGL(kYeeAfter<OPEN>, dim3(zt), dim3(256), 0, s, a, Enew[0], Enew[1], Enew[2], zc);
    if (absorb) return launchAbsorb(a.g, a.B, *ab, s);
    return hipGetLastError();
}

hipError_t launchFieldStep(int solver, const FieldArgs& a, float* const Enew[3], const AbsorbPlan* ab, hipStream_t s) {
    return a.g.open ? fieldStepT<true>(solver, a, Enew, ab, s) : fieldStepT<false>(solver, a, Enew, ab, s);
}

int fieldStepTiles(int solver, const GridDev& g) {
    if (solver == PIC_SOLVER_LEHE) return (g.nx / kFTX) * (g.ny / kFTY) * (g.nz / kFTZ);
    return zTileCount(g, kZC);
}

hipError_t launchFieldStepBatch(int solver, const FieldBatchEntry* table, const int2* tiles, int ntiles,
                                const int2* blocks, int nblocks, hipStream_t s) {
    (void)blocks;
    (void)nblocks;
    if (ntiles <= 0) return hipSuccess;
    if (solver == PIC_SOLVER_LEHE) {
        hipLaunchKernelGGL(kFieldBeforeAddBatch<PIC_SOLVER_LEHE>, dim3(ntiles), dim3(256), 0, s, table, tiles);
        hipLaunchKernelGGL(kFieldAfterBatch<PIC_SOLVER_LEHE>, dim3(ntiles), dim3(256), 0, s, table, tiles);
    } else {
        hipLaunchKernelGGL(kYeeBeforeAddBatch, dim3(ntiles), dim3(256), 0, s, table, tiles);
        hipLaunchKernelGGL(kYeeAfterBatch, dim3(ntiles), dim3(256), 0, s, table, tiles);
    }
    return hipGetLastError();
}

hipError_t launchCurrentResetBatch(const FieldBatchEntry* table, const int2* blocks, int nblocks, hipStream_t s) {
    if (nblocks <= 0) return hipSuccess;
    hipLaunchKernelGGL(kCurrentResetBatch, dim3(nblocks), dim3(256), 0, s, table, blocks);
    return hipGetLastError();
}

hipError_t launchAddCurrent(const FieldArgs& a, bool binomial, hipStream_t s) {
    size_t const n = (size_t)a.g.nx * a.g.ny * a.g.nz;
    if (binomial)
        hipLaunchKernelGGL(kAddCurrentBinomial, dim3(blocksFor(n)), dim3(256), 0, s, a);
    else
        hipLaunchKernelGGL(kAddCurrent, dim3(blocksFor(n)), dim3(256), 0, s, a);
    return hipGetLastError();
}

hipError_t launchBinomialJ(const FieldArgs& a, float* const out[3], hipStream_t s) {
    size_t const n = (size_t)a.g.nx * a.g.ny * a.g.nz;
    hipLaunchKernelGGL(kBinomialJ, dim3(blocksFor(n)), dim3(256), 0, s, a, out[0], out[1], out[2]);
    return hipGetLastError();
}

hipError_t launchCurrentReset(const GridDev& g, float* const J[3], size_t n, hipStream_t s) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(kCurrentReset, dim3(blocksFor(n)), dim3(256), 0, s, g, J[0], J[1], J[2], n);
    return hipGetLastError();
}

hipError_t launchEndStepBatch(const FieldBatchEntry* table, int n, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(kEndStepBatch, dim3((n + 63) / 64), dim3(64), 0, s, table, n);
    return hipGetLastError();
}

hipError_t launchFill(float* p, size_t n, float v, hipStream_t s) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(kFill, dim3(blocksFor(n)), dim3(256), 0, s, p, n, v);
    return hipGetLastError();
}

hipError_t launchEnergy(const FieldArgs& a, double* partials, int nblocks, double* out, hipStream_t s) {
    if (nblocks <= 0) return hipErrorInvalidValue;
    hipLaunchKernelGGL(kEnergy, dim3(nblocks), dim3(256), 0, s, a, partials);
    hipLaunchKernelGGL(kEnergyFinal, dim3(1), dim3(64), 0, s, partials, nblocks, out);
    return hipGetLastError();
}

hipError_t launchCountSum(const uint32_t* counts, int n, unsigned long long* out, hipStream_t s) {
    hipLaunchKernelGGL(kCountSum, dim3(1), dim3(256), 0, s, counts, n, out);
    return hipGetLastError();
}

}  // namespace pic
