// Differentiable form of the PIC step and its hand-written adjoint.
//
// Forward ("differentiable mode"): one mobile CIC species with the Boris
// pusher and the Yee solver, particles kept in identity order with global
// cell indices (no supercell shift), the same FP32 per-particle arithmetic as
// kPushDeposit (gather, Boris, moveParticle, Esirkepov), current accumulated
// as int64 fixed point in global memory (order independent), then the FDTD
// kernels of kernels_field.hip.
//
// Adjoint (FP64): reverse of one step given the adjoints of E, B at step n+1
// and of every particle's position X = cell + pos and momentum p:
//   fields:  transposed curls, (curl_forward)^T = curl_backward and vice versa;
//   current: lambda_J = -(dt/eps0) lambda_E;
//   particles: reverse Esirkepov written as a gather of lambda_J through the
//            derivatives of the CIC weights, reverse velocity, reverse Boris,
//            reverse field gather (scatter of lambda_E, lambda_B with FP64
//            atomics, derivative of the interpolation weights).
#include "pic_device.hpp"
#include "pic_kernels.hpp"

namespace pic {

__device__ __forceinline__ int wrapI(int v, int n) {
    v %= n;
    return v < 0 ? v + n : v;
}

__device__ __forceinline__ size_t gidx(const GridDev& g, int x, int y, int z) {
    return (size_t)wrapI(x, g.nx) + (size_t)g.nx * ((size_t)wrapI(y, g.ny) + (size_t)g.ny * (size_t)wrapI(z, g.nz));
}

// ---------------------------------------------------------------------------
// forward
// ---------------------------------------------------------------------------
// FieldToParticleInterpolation<CIC> on global memory: same arithmetic as gather1<2>.
__device__ __forceinline__ float gatherGlobalCIC(const float* __restrict__ F, const GridDev& g, int cx, int cy, int cz,
                                                 Vec3 pos, float fx, float fy, float fz) {
    float p[3] = {pos.x, pos.y, pos.z};
    float const f[3] = {fx, fy, fz};
    int is[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        float const v_pos = p[d] - f[d] - 0.5f;
        is[d] = v_pos >= -0.5f ? 0 : -1;
        p[d] = v_pos - (float)is[d] + 0.5f;
    }
    float sx[2], sy[2], sz[2];
    Shape<2>::onSupport(p[0], sx);
    Shape<2>::onSupport(p[1], sy);
    Shape<2>::onSupport(p[2], sz);
    float result_z = 0.0f;
#pragma unroll
    for (int z = 0; z < 2; ++z) {
        float result_y = 0.0f;
#pragma unroll
        for (int y = 0; y < 2; ++y) {
            float result_x = 0.0f;
#pragma unroll
            for (int x = 0; x < 2; ++x)
                result_x += F[gidx(g, cx + is[0] + x, cy + is[1] + y, cz + is[2] + z)] * sx[x];
            result_y += result_x * sy[y];
        }
        result_z += result_y * sz[z];
    }
    return result_z;
}



// Esirkepov CIC deposit of one particle into global int64 accumulators.
// Arithmetic identical to Esirkepov<2>::deposit; addresses wrap periodically.
__device__ void depositGlobalCIC(const DiffArgs& a, int cx, int cy, int cz, Vec3 pos, Vec3 vel, float charge) {
    constexpr int SUPP = 2;
    constexpr int begin = 0, end = 2;
    float const np[3] = {pos.x, pos.y, pos.z};
    float const dp[3] = {vel.x * a.dt / a.cell[0], vel.y * a.dt / a.cell[1], vel.z * a.dt / a.cell[2]};
    float p0[3], p1[3];
    int gs[3], st[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        p0[d] = np[d] - dp[d];
        p1[d] = np[d];
        int const iS = (int)floorf(p0[d]);
        int const iE = (int)floorf(p1[d]);
        gs[d] = iS < iE ? iS : iE;
        st[d] = (gs[d] == iS ? 2 : 0) | (gs[d] == iE ? 4 : 0) | (iS != iE ? 1 : 0);
        p0[d] -= (float)gs[d];
        p1[d] -= (float)gs[d];
    }
    int const bx = cx + gs[0], by = cy + gs[1], bz = cz + gs[2];
    // rotations as Esirkepov.hpp:106-124; perm maps rotated (i,j,k) to (x,y,z)
    for (int comp = 0; comp < 3; ++comp) {
        int ax[3];  // rotated axis -> real axis
        if (comp == 0) {
            ax[0] = 1; ax[1] = 2; ax[2] = 0;
        } else if (comp == 1) {
            ax[0] = 2; ax[1] = 0; ax[2] = 1;
        } else {
            ax[0] = 0; ax[1] = 1; ax[2] = 2;
        }
        float const q0[3] = {p0[ax[0]], p0[ax[1]], p0[ax[2]]};
        float const q1[3] = {p1[ax[0]], p1[ax[1]], p1[ax[2]]};
        int const s3[3] = {st[ax[0]], st[ax[1]], st[ax[2]]};
        if (q0[2] == q1[2]) continue;
        float const csd = charge * a.inv_vdt * a.cell[comp];
        float s0j[SUPP + 1], s1j[SUPP + 1], s0k[SUPP + 1], s1k[SUPP + 1];
        cachedShape<SUPP>(q0[1], !(s3[1] & 2), s0j);
        cachedShape<SUPP>(q1[1], !(s3[1] & 4), s1j);
        cachedShape<SUPP>(q0[2], !(s3[2] & 2), s0k);
        cachedShape<SUPP>(q1[2], !(s3[2] & 4), s1k);
        int const leaveI = s3[0] & 1, leaveJ = s3[1] & 1, leaveK = s3[2] & 1;
        for (int i = begin; i < end + 1; ++i) {
            if (!(i < end + leaveI)) continue;
            float const s0i = Shape<SUPP>::assign((float)i - q0[0]);
            float const dsi = Shape<SUPP>::assign((float)i - q1[0]) - s0i;
            for (int j = begin; j < end + 1; ++j) {
                if (!(j < end + leaveJ)) continue;
                float const sj0 = s0j[j - begin];
                float const dsj = s1j[j - begin] - sj0;
                float const tmp = -csd * (s0i * sj0 + 0.5f * (dsi * sj0 + s0i * dsj) + (1.0f / 3.0f) * dsj * dsi);
                float acc = 0.0f;
                for (int k = begin; k < end; ++k) {
                    if (!(k < end + leaveK - 1)) continue;
                    float const W = (s1k[k - begin] - s0k[k - begin]) * tmp;
                    acc += W;
                    int rel[3];
                    rel[ax[0]] = i;
                    rel[ax[1]] = j;
                    rel[ax[2]] = k;
                    atomicAdd(a.Jfix[comp] + gidx(a.g, bx + rel[0], by + rel[1], bz + rel[2]), toFixed(acc, a.jscale));
                }
            }
        }
    }
}

__global__ __launch_bounds__(256) void kDiffPush(DiffArgs a, DiffParticles P) {
    int64_t const i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= P.n) return;
    int64_t const n = P.n;
    int const cx = P.cell[i], cy = P.cell[n + i], cz = P.cell[2 * n + i];
    Vec3 const pos{P.pos[i], P.pos[n + i], P.pos[2 * n + i]};
    Vec3 mom{P.mom[i], P.mom[n + i], P.mom[2 * n + i]};
    float const w = P.w[i];
    float const mass = a.mass * w;
    float const charge = a.charge * w;
    Vec3 const e{gatherGlobalCIC(a.E[0], a.g, cx, cy, cz, pos, 0.5f, 0.0f, 0.0f),
                 gatherGlobalCIC(a.E[1], a.g, cx, cy, cz, pos, 0.0f, 0.5f, 0.0f),
                 gatherGlobalCIC(a.E[2], a.g, cx, cy, cz, pos, 0.0f, 0.0f, 0.5f)};
    Vec3 const b{gatherGlobalCIC(a.B[0], a.g, cx, cy, cz, pos, 0.0f, 0.5f, 0.5f),
                 gatherGlobalCIC(a.B[1], a.g, cx, cy, cz, pos, 0.5f, 0.0f, 0.5f),
                 gatherGlobalCIC(a.B[2], a.g, cx, cy, cz, pos, 0.5f, 0.5f, 0.0f)};
    PushConsts pk;
    for (int d = 0; d < 3; ++d) pk.cell[d] = a.cell[d];
    pk.dt = a.dt;
    pk.c2 = a.c2;
    pk.rc2 = a.rc2;
    pk.c = a.c;
    Vec3 vel;
    pushBoris(e, b, mom, vel, mass, charge, pk);
    float p[3] = {pos.x, pos.y, pos.z};
    p[0] += (vel.x * a.dt) / a.cell[0];
    p[1] += (vel.y * a.dt) / a.cell[1];
    p[2] += (vel.z * a.dt) / a.cell[2];
    int dir[3], idxUnused, maskUnused;
    moveParticle(p, 0, dir, idxUnused, maskUnused);
    int const nx = wrapI(cx + dir[0], a.g.nx), ny = wrapI(cy + dir[1], a.g.ny), nz = wrapI(cz + dir[2], a.g.nz);
    P.cell[i] = nx;
    P.cell[n + i] = ny;
    P.cell[2 * n + i] = nz;
    P.pos[i] = p[0];
    P.pos[n + i] = p[1];
    P.pos[2 * n + i] = p[2];
    P.mom[i] = mom.x;
    P.mom[n + i] = mom.y;
    P.mom[2 * n + i] = mom.z;
    depositGlobalCIC(a, nx, ny, nz, Vec3{p[0], p[1], p[2]}, vel, charge);
}

__global__ __launch_bounds__(256) void kFixToFloat(unsigned long long* fix, float* J, size_t n, double jinv) {
    size_t const i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    long long const acc = (long long)fix[i];
    J[i] = acc == 0 ? 0.0f : (float)((double)acc * jinv);
    fix[i] = 0ull;
}

// ---------------------------------------------------------------------------
// adjoint
// ---------------------------------------------------------------------------

// dst += coef * curl(src): forward differences (fwd = 1) or backward (fwd = 0)
template <int FWD>
__global__ __launch_bounds__(256) void kCurlAdd(AdjFieldArgs a, double* d0, double* d1, double* d2, const double* s0,
                                                const double* s1, const double* s2, double coef) {
    size_t const n = (size_t)a.g.nx * a.g.ny * a.g.nz;
    size_t const i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    int const x = (int)(i % a.g.nx), y = (int)((i / a.g.nx) % a.g.ny), z = (int)(i / ((size_t)a.g.nx * a.g.ny));
    const double* s[3] = {s0, s1, s2};
    auto der = [&](int c, int d) {
        int const ux = d == 0, uy = d == 1, uz = d == 2;
        if (FWD) return (s[c][gidx(a.g, x + ux, y + uy, z + uz)] - s[c][i]) / a.cell[d];
        return (s[c][i] - s[c][gidx(a.g, x - ux, y - uy, z - uz)]) / a.cell[d];
    };
    double const cx = der(2, 1) - der(1, 2);
    double const cy = der(0, 2) - der(2, 0);
    double const cz = der(1, 0) - der(0, 1);
    d0[i] += coef * cx;
    d1[i] += coef * cy;
    d2[i] += coef * cz;
}

__global__ __launch_bounds__(256) void kScaleCopy(const double* src, double* dst, size_t n, double coef) {
    size_t const i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) dst[i] = coef * src[i];
}



struct D3 {
    double x, y, z;
    __device__ double& operator[](int i) { return i == 0 ? x : (i == 1 ? y : z); }
    __device__ double operator[](int i) const { return i == 0 ? x : (i == 1 ? y : z); }
};
__device__ __forceinline__ D3 crossD(D3 a, D3 b) {
    return D3{a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
__device__ __forceinline__ double dotD(D3 a, D3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }

// CIC weight of node i for in-cell offset p (B-spline 1-|i-p|) and its derivative d/dp
__device__ __forceinline__ void cicW(double i, double p, double& w, double& dw) {
    double const u = i - p;
    double const au = fabs(u);
    if (au < 1.0) {
        w = 1.0 - au;
        dw = u > 0.0 ? 1.0 : (u < 0.0 ? -1.0 : 0.0);  // d(1-|i-p|)/dp = sign(i-p)
    } else {
        w = 0.0;
        dw = 0.0;
    }
}

__global__ __launch_bounds__(256) void kAdjParticles(AdjPartArgs a, AdjState S) {
    int64_t const idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= S.n) return;
    int64_t const n = S.n;
    int const c[3] = {S.cell[idx], S.cell[n + idx], S.cell[2 * n + idx]};
    double const r[3] = {(double)S.pos[idx], (double)S.pos[n + idx], (double)S.pos[2 * n + idx]};
    D3 const p0{(double)S.mom[idx], (double)S.mom[n + idx], (double)S.mom[2 * n + idx]};
    double const w = (double)S.w[idx];
    double const m = a.mass * w;
    double const q = a.charge * w;
    double const dt = a.dt;

    // ---- forward recompute in FP64: gather (weights and their derivatives) ----
    const double fE[3][3] = {{0.5, 0, 0}, {0, 0.5, 0}, {0, 0, 0.5}};
    const double fB[3][3] = {{0, 0.5, 0.5}, {0.5, 0, 0.5}, {0.5, 0.5, 0}};
    int baseE[3][3], baseB[3][3];
    double wE[3][3][2], dwE[3][3][2], wB[3][3][2], dwB[3][3][2];
    D3 Ep{0, 0, 0}, Bp{0, 0, 0};
    D3 dEp[3], dBp[3];  // dEp[comp][axis]
    for (int comp = 0; comp < 3; ++comp) {
        for (int d = 0; d < 3; ++d) {
            double const v = r[d] - fE[comp][d] - 0.5;
            int const s = v >= -0.5 ? 0 : -1;
            double const pp = v - s + 0.5;
            baseE[comp][d] = c[d] + s;
            wE[comp][d][0] = 1.0 - pp;
            wE[comp][d][1] = pp;
            dwE[comp][d][0] = -1.0;
            dwE[comp][d][1] = 1.0;
            double const vb = r[d] - fB[comp][d] - 0.5;
            int const sb = vb >= -0.5 ? 0 : -1;
            double const pb = vb - sb + 0.5;
            baseB[comp][d] = c[d] + sb;
            wB[comp][d][0] = 1.0 - pb;
            wB[comp][d][1] = pb;
            dwB[comp][d][0] = -1.0;
            dwB[comp][d][1] = 1.0;
        }
        double ve = 0, vb = 0;
        D3 ge{0, 0, 0}, gb{0, 0, 0};
        for (int zz = 0; zz < 2; ++zz)
            for (int yy = 0; yy < 2; ++yy)
                for (int xx = 0; xx < 2; ++xx) {
                    double const fe = a.E[comp][gidx(a.g, baseE[comp][0] + xx, baseE[comp][1] + yy, baseE[comp][2] + zz)];
                    double const fb = a.B[comp][gidx(a.g, baseB[comp][0] + xx, baseB[comp][1] + yy, baseB[comp][2] + zz)];
                    double const we = wE[comp][0][xx] * wE[comp][1][yy] * wE[comp][2][zz];
                    double const wb = wB[comp][0][xx] * wB[comp][1][yy] * wB[comp][2][zz];
                    ve += fe * we;
                    vb += fb * wb;
                    ge.x += fe * dwE[comp][0][xx] * wE[comp][1][yy] * wE[comp][2][zz];
                    ge.y += fe * wE[comp][0][xx] * dwE[comp][1][yy] * wE[comp][2][zz];
                    ge.z += fe * wE[comp][0][xx] * wE[comp][1][yy] * dwE[comp][2][zz];
                    gb.x += fb * dwB[comp][0][xx] * wB[comp][1][yy] * wB[comp][2][zz];
                    gb.y += fb * wB[comp][0][xx] * dwB[comp][1][yy] * wB[comp][2][zz];
                    gb.z += fb * wB[comp][0][xx] * wB[comp][1][yy] * dwB[comp][2][zz];
                }
        Ep[comp] = ve;
        Bp[comp] = vb;
        dEp[comp] = ge;
        dBp[comp] = gb;
    }
    // ---- Boris forward (particlePusherBoris.hpp:41-90) ----
    double const hq = 0.5 * q * dt;
    D3 const pm{p0.x + hq * Ep.x, p0.y + hq * Ep.y, p0.z + hq * Ep.z};
    double const g = sqrt(1.0 + dotD(pm, pm) * a.rc2 / (m * m));
    double const beta = 0.5 * (q / m) * dt;
    D3 const t{beta * Bp.x / g, beta * Bp.y / g, beta * Bp.z / g};
    double const f = 1.0 / (1.0 + dotD(t, t));
    D3 const s{2.0 * f * t.x, 2.0 * f * t.y, 2.0 * f * t.z};
    D3 const cpt = crossD(pm, t);
    D3 const pp{pm.x + cpt.x, pm.y + cpt.y, pm.z + cpt.z};
    D3 const cps = crossD(pp, s);
    D3 const p1{pm.x + cps.x + hq * Ep.x, pm.y + cps.y + hq * Ep.y, pm.z + cps.z + hq * Ep.z};
    double const G = sqrt(m * m + dotD(p1, p1) * a.rc2);
    D3 const v{p1.x / G, p1.y / G, p1.z / G};
    // end position in cell units relative to the particle's cell: r + v dt / cell
    double X1[3], X0[3];
    for (int d = 0; d < 3; ++d) {
        X1[d] = r[d] + v[d] * dt / a.cell[d];
        X0[d] = X1[d] - v[d] * dt / a.cell[d];
    }

    // ---- reverse Esirkepov as a gather of lambda_J ----
    // shape values at nodes gs+0..2 per axis, for start (0) and end (1)
    int gs[3];
    double s0[3][3], s1[3][3], ds0[3][3], ds1[3][3];
    for (int d = 0; d < 3; ++d) {
        gs[d] = (int)fmin(floor(X0[d]), floor(X1[d]));
        for (int k = 0; k < 3; ++k) {
            cicW((double)k, X0[d] - gs[d], s0[d][k], ds0[d][k]);
            cicW((double)k, X1[d] - gs[d], s1[d][k], ds1[d][k]);
        }
    }
    double g0[3][3] = {{0}}, g1[3][3] = {{0}};  // dLambda/d s0, s1 per axis and node
    for (int comp = 0; comp < 3; ++comp) {
        int const ax = comp == 0 ? 1 : (comp == 1 ? 2 : 0);  // rotated i axis
        int const bx = comp == 0 ? 2 : (comp == 1 ? 0 : 1);  // rotated j axis
        double const csd = q * a.inv_vdt * a.cell[comp];
        double C[3], lam[3][3][3];
        double run = 0.0;
        for (int k = 0; k < 3; ++k) {
            run += s1[comp][k] - s0[comp][k];
            C[k] = run;
        }
        double dC[3] = {0, 0, 0};
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j) {
                double const s0a = s0[ax][i], dsa = s1[ax][i] - s0[ax][i];
                double const s0b = s0[bx][j], dsb = s1[bx][j] - s0[bx][j];
                double const W = s0a * s0b + 0.5 * (dsa * s0b + s0a * dsb) + (1.0 / 3.0) * dsa * dsb;
                double Gij = 0.0;
                for (int k = 0; k < 2; ++k) {  // the current at the last node of the axis is zero
                    int rel[3];
                    rel[ax] = i;
                    rel[bx] = j;
                    rel[comp] = k;
                    double const l = a.lJ[comp][gidx(a.g, c[0] + gs[0] + rel[0], c[1] + gs[1] + rel[1],
                                                     c[2] + gs[2] + rel[2])];
                    lam[i][j][k] = l;
                    Gij += -csd * l * C[k];
                    dC[k] += -csd * l * W;
                }
                double const dWs0a = s0b + 0.5 * dsb, dWdsa = 0.5 * s0b + (1.0 / 3.0) * dsb;
                double const dWs0b = s0a + 0.5 * dsa, dWdsb = 0.5 * s0a + (1.0 / 3.0) * dsa;
                g0[ax][i] += Gij * (dWs0a - dWdsa);
                g1[ax][i] += Gij * dWdsa;
                g0[bx][j] += Gij * (dWs0b - dWdsb);
                g1[bx][j] += Gij * dWdsb;
            }
        // C[k] = sum_{k'<=k} (s1 - s0): reverse prefix sum
        double tail = 0.0;
        for (int k = 2; k >= 0; --k) {
            tail += dC[k];
            g1[comp][k] += tail;
            g0[comp][k] -= tail;
        }
    }
    // weights depend on X0 = X1 - v dt/cell and X1 (offsets relative to gs)
    double lX1[3], lv[3];
    for (int d = 0; d < 3; ++d) {
        double l0 = 0.0, l1 = 0.0;
        for (int k = 0; k < 3; ++k) {
            l0 += g0[d][k] * ds0[d][k];
            l1 += g1[d][k] * ds1[d][k];
        }
        lX1[d] = S.lX[d * n + idx] + l0 + l1;
        lv[d] = -l0 * dt / a.cell[d];
    }
    // X1 = X + v dt/cell
    for (int d = 0; d < 3; ++d) lv[d] += lX1[d] * dt / a.cell[d];
    // v = p1 / G, G = sqrt(m^2 + |p1|^2 rc2)
    D3 const LV{lv[0], lv[1], lv[2]};
    double const pv = dotD(p1, LV);
    D3 l1{S.lP[idx] + LV.x / G - p1.x * pv * a.rc2 / (G * G * G),
          S.lP[n + idx] + LV.y / G - p1.y * pv * a.rc2 / (G * G * G),
          S.lP[2 * n + idx] + LV.z / G - p1.z * pv * a.rc2 / (G * G * G)};
    // ---- reverse Boris ----
    D3 lEp{hq * l1.x, hq * l1.y, hq * l1.z};
    D3 lpm = l1;                 // pl = pm + pp x s
    D3 const lpp = crossD(s, l1);   // d(pp x s)/dpp ^T l = s x l
    D3 const ls = crossD(l1, pp);   // d(pp x s)/ds ^T l = l x pp
    // pp = pm + pm x t
    D3 const tl = crossD(t, lpp);
    lpm = D3{lpm.x + lpp.x + tl.x, lpm.y + lpp.y + tl.y, lpm.z + lpp.z + tl.z};
    D3 const lt1 = crossD(lpp, pm);
    // s = 2 f t, f = 1/(1+|t|^2)
    double const tls = dotD(t, ls);
    D3 const lt{lt1.x + 2.0 * f * ls.x - 4.0 * f * f * tls * t.x, lt1.y + 2.0 * f * ls.y - 4.0 * f * f * tls * t.y,
                lt1.z + 2.0 * f * ls.z - 4.0 * f * f * tls * t.z};
    // t = beta B / g
    D3 const lBp{beta * lt.x / g, beta * lt.y / g, beta * lt.z / g};
    double const lg = -beta * dotD(Bp, lt) / (g * g);
    // g = sqrt(1 + |pm|^2 rc2/m^2)
    double const kk = lg * a.rc2 / (m * m * g);
    lpm = D3{lpm.x + kk * pm.x, lpm.y + kk * pm.y, lpm.z + kk * pm.z};
    // pm = p0 + hq E
    lEp = D3{lEp.x + hq * lpm.x, lEp.y + hq * lpm.y, lEp.z + hq * lpm.z};
    D3 const lp0 = lpm;
    // ---- reverse gather ----
    double lX0[3] = {lX1[0], lX1[1], lX1[2]};
    for (int comp = 0; comp < 3; ++comp) {
        lX0[0] += lEp[comp] * dEp[comp].x + lBp[comp] * dBp[comp].x;
        lX0[1] += lEp[comp] * dEp[comp].y + lBp[comp] * dBp[comp].y;
        lX0[2] += lEp[comp] * dEp[comp].z + lBp[comp] * dBp[comp].z;
        for (int zz = 0; zz < 2; ++zz)
            for (int yy = 0; yy < 2; ++yy)
                for (int xx = 0; xx < 2; ++xx) {
                    double const we = wE[comp][0][xx] * wE[comp][1][yy] * wE[comp][2][zz];
                    double const wb = wB[comp][0][xx] * wB[comp][1][yy] * wB[comp][2][zz];
                    atomicAdd(a.lE[comp] + gidx(a.g, baseE[comp][0] + xx, baseE[comp][1] + yy, baseE[comp][2] + zz),
                              lEp[comp] * we);
                    atomicAdd(a.lB[comp] + gidx(a.g, baseB[comp][0] + xx, baseB[comp][1] + yy, baseB[comp][2] + zz),
                              lBp[comp] * wb);
                }
    }
    for (int d = 0; d < 3; ++d) {
        S.lX[d * n + idx] = lX0[d];
        S.lP[d * n + idx] = lp0[d];
    }
}

__global__ __launch_bounds__(256) void kFillD(double* p, size_t n, double v) {
    size_t const i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------
static inline unsigned blocksOf(size_t n) { return (unsigned)((n + 255) / 256); }

hipError_t launchDiffPush(const DiffArgs& a, const DiffParticles& P, hipStream_t s) {
    if (P.n == 0) return hipSuccess;
    hipLaunchKernelGGL(kDiffPush, dim3(blocksOf((size_t)P.n)), dim3(256), 0, s, a, P);
    return hipGetLastError();
}

hipError_t launchFixToFloat(unsigned long long* fix, float* J, size_t n, double jinv, hipStream_t s) {
    hipLaunchKernelGGL(kFixToFloat, dim3(blocksOf(n)), dim3(256), 0, s, fix, J, n, jinv);
    return hipGetLastError();
}

// reverse of: B' = B - h C+ E; E' = E + dt c^2 C- B' + coeff J; B'' = B' - h C+ E'
// given lE, lB (adjoints of E'', B''), leaves the adjoints of E, B there and
// lambda_J = coeff * (adjoint of E') in lJ.
hipError_t launchAdjFields(const AdjFieldArgs& a, double dt, double c2, double coeff, double* lJ[3], hipStream_t s) {
    size_t const n = (size_t)a.g.nx * a.g.ny * a.g.nz;
    double const h = 0.5 * dt;
    // B'' = B' - h C+ E'  ->  lE' = lE - h C- lB
    hipLaunchKernelGGL(kCurlAdd<0>, dim3(blocksOf(n)), dim3(256), 0, s, a, a.lE[0], a.lE[1], a.lE[2], a.lB[0],
                       a.lB[1], a.lB[2], -h);
    for (int c = 0; c < 3; ++c)
        hipLaunchKernelGGL(kScaleCopy, dim3(blocksOf(n)), dim3(256), 0, s, a.lE[c], lJ[c], n, coeff);
    // E' = E + dt c^2 C- B' + coeff J  ->  lB' = lB + dt c^2 C+ lE'
    hipLaunchKernelGGL(kCurlAdd<1>, dim3(blocksOf(n)), dim3(256), 0, s, a, a.lB[0], a.lB[1], a.lB[2], a.lE[0],
                       a.lE[1], a.lE[2], dt * c2);
    // B' = B - h C+ E  ->  lE = lE' - h C- lB'
    hipLaunchKernelGGL(kCurlAdd<0>, dim3(blocksOf(n)), dim3(256), 0, s, a, a.lE[0], a.lE[1], a.lE[2], a.lB[0],
                       a.lB[1], a.lB[2], -h);
    return hipGetLastError();
}

hipError_t launchFillD(double* p, size_t n, double v, hipStream_t s) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(kFillD, dim3(blocksOf(n)), dim3(256), 0, s, p, n, v);
    return hipGetLastError();
}

hipError_t launchAdjParticles(const AdjPartArgs& a, const AdjState& st, hipStream_t s) {
    if (st.n == 0) return hipSuccess;
    hipLaunchKernelGGL(kAdjParticles, dim3(blocksOf((size_t)st.n)), dim3(256), 0, s, a, st);
    return hipGetLastError();
}

}  // namespace pic
