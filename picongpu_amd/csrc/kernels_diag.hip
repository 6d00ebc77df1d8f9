// Diagnostics of the PIC step on the device.
//
// Gauss-law residual eps0 * div E - rho at the cell origins in FP64, the
// quantity ChargeConservation::notify reports
// (plugins/ChargeConservation.tpp:181-260: backward-difference divergence
// `Div` :118-136, charge density from the particles' assignment shape). The
// charge density uses each species' ChargeAssignment B-spline evaluated in
// FP32 per axis and multiplied in FP64; it is accumulated with FP64 atomics,
// so the residual is reproducible to FP64 round-off (not bit for bit).
#include <array>
#include <mutex>

#include "pic_device.hpp"
#include "pic_kernels.hpp"

namespace pic {

template <int SUPP>
__global__ __launch_bounds__(256) void kChargeDensity(SpeciesDev S, GridDev g, double qScale, double* rho) {
    int const sc = blockIdx.x;
    int const sx = sc % g.nscx;
    int const sy = (sc / g.nscx) % g.nscy;
    int const sz = sc / (g.nscx * g.nscy);
    int const n = (int)S.count[sc];
    size_t const base = slotBase(S, sc);
    constexpr int begin = Shape<SUPP>::begin;
    for (int i = threadIdx.x; i < n; i += blockDim.x) {
        size_t const gi = base + i;
        int const lc = S.cell[gi];
        int const c[3] = {sx * SCX + lc % SCX, sy * SCY + (lc / SCX) % SCY, sz * SCZ + lc / (SCX * SCY)};
        float const p[3] = {S.pos[0][gi], S.pos[1][gi], S.pos[2][gi]};
        double const q = (double)S.charge * (double)S.w[gi] * qScale;
        double wgt[3][SUPP + 1];
#pragma unroll
        for (int d = 0; d < 3; ++d)
#pragma unroll
            for (int o = 0; o <= SUPP; ++o) wgt[d][o] = (double)Shape<SUPP>::assign((float)(begin + o) - p[d]);
        for (int oz = 0; oz <= SUPP; ++oz) {
            int zz = c[2] + begin + oz;
            zz = zz < 0 ? zz + g.nz : (zz >= g.nz ? zz - g.nz : zz);
            for (int oy = 0; oy <= SUPP; ++oy) {
                int yy = c[1] + begin + oy;
                yy = yy < 0 ? yy + g.ny : (yy >= g.ny ? yy - g.ny : yy);
                for (int ox = 0; ox <= SUPP; ++ox) {
                    double const v = q * wgt[0][ox] * wgt[1][oy] * wgt[2][oz];
                    if (v == 0.0 || outsideOpen(g, c[0] + begin + ox, c[1] + begin + oy, c[2] + begin + oz)) continue;
                    int xx = c[0] + begin + ox;
                    xx = xx < 0 ? xx + g.nx : (xx >= g.nx ? xx - g.nx : xx);
                    unsafeAtomicAdd(rho + (size_t)xx + (size_t)g.nx * ((size_t)yy + (size_t)g.ny * zz), v);
                }
            }
        }
    }
}

__global__ __launch_bounds__(256) void kGaussResidual(FieldArgs a, double eps0, double* inout) {
    size_t const n = (size_t)a.g.nx * a.g.ny * a.g.nz;
    size_t const i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    int const x = (int)(i % a.g.nx);
    int const y = (int)((i / a.g.nx) % a.g.ny);
    int const z = (int)(i / ((size_t)a.g.nx * a.g.ny));
    int const xm = x == 0 ? a.g.nx - 1 : x - 1;
    int const ym = y == 0 ? a.g.ny - 1 : y - 1;
    int const zm = z == 0 ? a.g.nz - 1 : z - 1;
    size_t const ix = (size_t)xm + (size_t)a.g.nx * ((size_t)y + (size_t)a.g.ny * z);
    size_t const iy = (size_t)x + (size_t)a.g.nx * ((size_t)ym + (size_t)a.g.ny * z);
    size_t const iz = (size_t)x + (size_t)a.g.nx * ((size_t)y + (size_t)a.g.ny * zm);
    // beyond an open border E reads zero
    double const ex = (a.g.open & 1) && x == 0 ? 0.0 : (double)a.E[0][ix];
    double const ey = (a.g.open & 2) && y == 0 ? 0.0 : (double)a.E[1][iy];
    double const ez = (a.g.open & 4) && z == 0 ? 0.0 : (double)a.E[2][iz];
    double div = ((double)a.E[0][i] - ex) / (double)a.cell[0];
    div += ((double)a.E[1][i] - ey) / (double)a.cell[1];
    div += ((double)a.E[2][i] - ez) / (double)a.cell[2];
    inout[i] = eps0 * div - inout[i];
}

__global__ __launch_bounds__(256) void kZeroD(double* p, size_t n) {
    size_t const i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0.0;
}

// Scratch of the allocation-free Gauss residual: one FP64 value per cell,
// at most kNoAllocMaxCells cells. Each device holds its own copy for the
// life of the process; host code reaches it only through the address
// hipGetSymbolAddress returns on that device.
static __device__ __align__(256) double scratch[kNoAllocMaxCells];

hipError_t noAllocScratch(int device, double** out) {
    static std::mutex lock;
    static std::array<double*, kNoAllocMaxDevices> addr{};
    if (device < 0 || device >= kNoAllocMaxDevices) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> guard(lock);
    if (!addr[device]) {
        int current = 0;
        hipError_t e = hipGetDevice(&current);
        if (e != hipSuccess) return e;
        if (current != device) return hipErrorInvalidDevice;  // the caller makes `device` current
        void* p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return e;
        addr[device] = static_cast<double*>(p);
    }
    *out = addr[device];
    return hipSuccess;
}

hipError_t launchGaussResidual(const FieldArgs& f, const SpeciesDev* sp, const int* supp, int nsp, double qScale,
                               double eps0, double* buf, hipStream_t s) {
    size_t const n = (size_t)f.g.nx * f.g.ny * f.g.nz;
    int const nsc = f.g.nscx * f.g.nscy * f.g.nscz;
    hipLaunchKernelGGL(kZeroD, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, buf, n);
    for (int k = 0; k < nsp; ++k) {
        if (supp[k] == 2)
            hipLaunchKernelGGL(kChargeDensity<2>, dim3(nsc), dim3(256), 0, s, sp[k], f.g, qScale, buf);
        else if (supp[k] == 3)
            hipLaunchKernelGGL(kChargeDensity<3>, dim3(nsc), dim3(256), 0, s, sp[k], f.g, qScale, buf);
        else
            hipLaunchKernelGGL(kChargeDensity<4>, dim3(nsc), dim3(256), 0, s, sp[k], f.g, qScale, buf);
        hipError_t const e = hipGetLastError();
        if (e != hipSuccess) return e;
    }
    hipLaunchKernelGGL(kGaussResidual, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, f, eps0, buf);
    return hipGetLastError();
}

}  // namespace pic
