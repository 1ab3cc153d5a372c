// 2DGS projection (ray-splat transform, screen bounds, normals) and its VJP.
//
// Semantics follow textured_gaussians/cuda/csrc/fully_fused_projection_2dgs_fwd.cu:72-238
// (forward) and utils.cuh:659-791 (reference VJPs); TGS_VJP_EXACT is the true
// derivative of the same forward.  One thread per (camera, Gaussian); the
// Gaussian index runs fastest so a wavefront reads 64 consecutive Gaussians.
#include "common.h"

namespace tgs {

template <typename T> struct Cam {
    m33<T> R;
    v3<T> t;
    T fx, fy, cx, cy;
};

template <typename T> __device__ __forceinline__ Cam<T> load_cam(const T *viewmats, const T *Ks, int64_t c) {
    Cam<T> cam;
    const T *v = viewmats + c * 16;
#pragma unroll
    for (int r = 0; r < 3; ++r)
#pragma unroll
        for (int k = 0; k < 3; ++k) cam.R.m[r][k] = v[r * 4 + k];
    cam.t = mk3(v[3], v[7], v[11]);
    const T *K = Ks + c * 9;
    cam.fx = K[0];
    cam.cx = K[2];
    cam.fy = K[4];
    cam.cy = K[5];
    return cam;
}

// Rotation matrix of the normalised (w,x,y,z) quaternion.
template <typename T> __device__ __forceinline__ m33<T> quat_rot(T w, T x, T y, T z) {
    const T inv = tgs_rsqrt<T>(w * w + x * x + y * y + z * z);
    w *= inv; x *= inv; y *= inv; z *= inv;
    m33<T> R;
    R.m[0][0] = T(1) - T(2) * (y * y + z * z);
    R.m[0][1] = T(2) * (x * y - w * z);
    R.m[0][2] = T(2) * (x * z + w * y);
    R.m[1][0] = T(2) * (x * y + w * z);
    R.m[1][1] = T(1) - T(2) * (x * x + z * z);
    R.m[1][2] = T(2) * (y * z - w * x);
    R.m[2][0] = T(2) * (x * z - w * y);
    R.m[2][1] = T(2) * (y * z + w * x);
    R.m[2][2] = T(1) - T(2) * (x * x + y * y);
    return R;
}

// d(loss)/d(quat) given d(loss)/d(R) for R = quat_rot(q) (normalisation included).
template <typename T> __device__ __forceinline__ void quat_rot_vjp(const T q[4], const m33<T> &G, T vq[4]) {
    const T inv = tgs_rsqrt<T>(q[0] * q[0] + q[1] * q[1] + q[2] * q[2] + q[3] * q[3]);
    const T w = q[0] * inv, x = q[1] * inv, y = q[2] * inv, z = q[3] * inv;
    // derivative of each entry of R w.r.t. the unit quaternion, contracted with G
    T gw = T(2) * (x * (G.m[2][1] - G.m[1][2]) + y * (G.m[0][2] - G.m[2][0]) + z * (G.m[1][0] - G.m[0][1]));
    T gx = T(2) * (-T(2) * x * (G.m[1][1] + G.m[2][2]) + y * (G.m[1][0] + G.m[0][1]) +
                   z * (G.m[2][0] + G.m[0][2]) + w * (G.m[2][1] - G.m[1][2]));
    T gy = T(2) * (x * (G.m[1][0] + G.m[0][1]) - T(2) * y * (G.m[0][0] + G.m[2][2]) +
                   z * (G.m[2][1] + G.m[1][2]) + w * (G.m[0][2] - G.m[2][0]));
    T gz = T(2) * (x * (G.m[2][0] + G.m[0][2]) + y * (G.m[2][1] + G.m[1][2]) -
                   T(2) * z * (G.m[0][0] + G.m[1][1]) + w * (G.m[1][0] - G.m[0][1]));
    // project out the radial component (normalisation) and scale by 1/|q|
    const T d = gw * w + gx * x + gy * y + gz * z;
    vq[0] = (gw - d * w) * inv;
    vq[1] = (gx - d * x) * inv;
    vq[2] = (gy - d * y) * inv;
    vq[3] = (gz - d * z) * inv;
}

// Projection of Gaussian g into camera c (fully_fused_projection_fwd_2dgs_kernel,
// projection_2dgs_fwd.cu:72-238): ok = visible and inside the image after the
// radius clip; then the outputs below are set.
template <typename T> struct Proj {
    bool ok;
    int32_t radius;
    T mx, my, depth;
    T rt[9];
    T n[3];
};

template <typename T>
__device__ __forceinline__ Proj<T> project_one(int64_t c, int64_t g, const T *__restrict__ means,
                                               const T *__restrict__ quats, const T *__restrict__ scales,
                                               const T *__restrict__ viewmats, const T *__restrict__ Ks, int32_t W,
                                               int32_t H, T near_plane, T far_plane, T radius_clip) {
    Proj<T> P;
    const Cam<T> cam = load_cam(viewmats, Ks, c);
    const v3<T> mu = mk3(means[g * 3], means[g * 3 + 1], means[g * 3 + 2]);
    const v3<T> pc = matvec(cam.R, mu) + cam.t;

    bool ok = !(pc.z < near_plane || pc.z > far_plane);
    const m33<T> Rq = quat_rot(quats[g * 4], quats[g * 4 + 1], quats[g * 4 + 2], quats[g * 4 + 3]);
    const T s0 = scales[g * 3], s1 = scales[g * 3 + 1];
    const m33<T> RR = matmul(cam.R, Rq);  // camera-frame rotation of the splat
    // columns of R Rq diag(s0, s1, 1)
    const v3<T> a0 = mk3(RR.m[0][0] * s0, RR.m[1][0] * s0, RR.m[2][0] * s0);
    const v3<T> a1 = mk3(RR.m[0][1] * s1, RR.m[1][1] * s1, RR.m[2][1] * s1);
    const v3<T> nz = mk3(RR.m[0][2], RR.m[1][2], RR.m[2][2]);
    // T = K [a0 a1 pc], rows T0, T1, T2
    const v3<T> T0 = mk3(cam.fx * a0.x + cam.cx * a0.z, cam.fx * a1.x + cam.cx * a1.z, cam.fx * pc.x + cam.cx * pc.z);
    const v3<T> T1 = mk3(cam.fy * a0.y + cam.cy * a0.z, cam.fy * a1.y + cam.cy * a1.z, cam.fy * pc.y + cam.cy * pc.z);
    const v3<T> T2 = mk3(a0.z, a1.z, pc.z);

    const T dist = T2.x * T2.x + T2.y * T2.y - T2.z * T2.z;
    ok = ok && dist != T(0);
    T mx = 0, my = 0, radius = 0;
    if (ok) {
        const T f = T(1) / dist;
        mx = f * (T0.x * T2.x + T0.y * T2.y - T0.z * T2.z);
        my = f * (T1.x * T2.x + T1.y * T2.y - T1.z * T2.z);
        const T hx = mx * mx - f * (T0.x * T0.x + T0.y * T0.y - T0.z * T0.z);
        const T hy = my * my - f * (T1.x * T1.x + T1.y * T1.y - T1.z * T1.z);
        radius = ceil(T(3) * sqrt(fmax(T(1e-4), fmax(hx, hy))));
        ok = radius > radius_clip && !(mx + radius <= T(0) || mx - radius >= T(W) || my + radius <= T(0) ||
                                       my - radius >= T(H));
    }
    P.ok = ok;
    P.radius = ok ? (int32_t)radius : 0;
    P.mx = ok ? mx : T(0);
    P.my = ok ? my : T(0);
    P.depth = ok ? pc.z : T(0);
    const T sgn = dot(nz, pc) < T(0) ? T(1) : T(-1);  // face the camera: -n.pc > 0
    const T tv[9] = {T0.x, T0.y, T0.z, T1.x, T1.y, T1.z, T2.x, T2.y, T2.z};
#pragma unroll
    for (int k = 0; k < 9; ++k) P.rt[k] = ok ? tv[k] : T(0);
    P.n[0] = ok ? sgn * nz.x : T(0);
    P.n[1] = ok ? sgn * nz.y : T(0);
    P.n[2] = ok ? sgn * nz.z : T(0);
    return P;
}

template <typename T>
__device__ __forceinline__ void store_proj(const Proj<T> &P, int64_t o, int32_t *__restrict__ radii,
                                           T *__restrict__ means2d, T *__restrict__ depths, T *__restrict__ rt,
                                           T *__restrict__ normals) {
    radii[o] = P.radius;
    means2d[o * 2] = P.mx;
    means2d[o * 2 + 1] = P.my;
    depths[o] = P.depth;
#pragma unroll
    for (int k = 0; k < 9; ++k) rt[o * 9 + k] = P.rt[k];
#pragma unroll
    for (int k = 0; k < 3; ++k) normals[o * 3 + k] = P.n[k];
}

template <typename T>
__global__ __launch_bounds__(256) void projection_fwd_kernel(
    int64_t C, int64_t N, const T *__restrict__ means, const T *__restrict__ quats,
    const T *__restrict__ scales, const T *__restrict__ viewmats, const T *__restrict__ Ks,
    int32_t W, int32_t H, T near_plane, T far_plane, T radius_clip, int32_t *__restrict__ radii,
    T *__restrict__ means2d, T *__restrict__ depths, T *__restrict__ rt, T *__restrict__ normals) {
    const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= C * N) return;
    const Proj<T> P = project_one(idx / N, idx % N, means, quats, scales, viewmats, Ks, W, H, near_plane, far_plane,
                                  radius_clip);
    store_proj(P, idx, radii, means2d, depths, rt, normals);
}

// Packed layout, pass 1: visibility flag per (camera, Gaussian), row-major.
template <typename T>
__global__ __launch_bounds__(256) void projection_flag_kernel(
    int64_t C, int64_t N, const T *__restrict__ means, const T *__restrict__ quats,
    const T *__restrict__ scales, const T *__restrict__ viewmats, const T *__restrict__ Ks,
    int32_t W, int32_t H, T near_plane, T far_plane, T radius_clip, int32_t *__restrict__ flags) {
    const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= C * N) return;
    flags[idx] = project_one(idx / N, idx % N, means, quats, scales, viewmats, Ks, W, H, near_plane, far_plane,
                             radius_clip).ok ? 1 : 0;
}

// Packed layout, pass 2: the visible entries at their rank among the visible ones
// (incl: inclusive prefix sum of the flags), in (camera, Gaussian) order; indptr[c]
// = first entry of camera c.
template <typename T>
__global__ __launch_bounds__(256) void projection_packed_kernel(
    int64_t C, int64_t N, const T *__restrict__ means, const T *__restrict__ quats,
    const T *__restrict__ scales, const T *__restrict__ viewmats, const T *__restrict__ Ks,
    int32_t W, int32_t H, T near_plane, T far_plane, T radius_clip, const int64_t *__restrict__ incl,
    int64_t *__restrict__ indptr, int64_t *__restrict__ camera_ids, int64_t *__restrict__ gaussian_ids,
    int32_t *__restrict__ radii, T *__restrict__ means2d, T *__restrict__ depths, T *__restrict__ rt,
    T *__restrict__ normals) {
    const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx <= C && indptr) indptr[idx] = idx == 0 ? 0 : incl[idx * N - 1];
    if (idx >= C * N) return;
    const int64_t before = idx == 0 ? 0 : incl[idx - 1];
    if (incl[idx] == before) return;  // not visible
    const int64_t c = idx / N, g = idx % N;
    const Proj<T> P = project_one(c, g, means, quats, scales, viewmats, Ks, W, H, near_plane, far_plane, radius_clip);
    camera_ids[before] = c;
    gaussian_ids[before] = g;
    store_proj(P, before, radii, means2d, depths, rt, normals);
}

// Backward over E entries: the C x N grid (camera_ids == NULL, entry c*N + g) or a
// packed list (entry i is camera camera_ids[i], Gaussian gaussian_ids[i], cameras
// in ascending order).  sparse: the Gaussian gradients are written per entry
// ([E, ...]) instead of accumulated per Gaussian ([N, ...]).
template <typename T>
__global__ __launch_bounds__(256) void projection_bwd_kernel(
    int64_t C, int64_t N, int64_t E, const int64_t *__restrict__ camera_ids, const int64_t *__restrict__ gaussian_ids,
    int sparse, const T *__restrict__ means, const T *__restrict__ quats,
    const T *__restrict__ scales, const T *__restrict__ viewmats, const T *__restrict__ Ks,
    const int32_t *__restrict__ radii, const T *__restrict__ rt, const T *__restrict__ v_means2d,
    const T *__restrict__ v_depths, const T *__restrict__ v_normals, const T *__restrict__ v_rt,
    int mode, T *__restrict__ v_means, T *__restrict__ v_quats, T *__restrict__ v_scales,
    T *__restrict__ v_viewmats) {
    const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const bool active = idx < E && (radii == nullptr || radii[idx] > 0);  // packed entries are all visible
    const int64_t c = active ? (camera_ids ? camera_ids[idx] : idx / N) : 0;
    const int64_t g = active ? (gaussian_ids ? gaussian_ids[idx] : idx % N) : 0;
    const int64_t og = sparse ? idx : g;  // where this entry's Gaussian gradients go
    T vR[3][4] = {};  // exact-mode d/d(viewmat[0:3, 0:4])
    if (active) {
        const Cam<T> cam = load_cam(viewmats, Ks, c);
        const v3<T> mu = mk3(means[g * 3], means[g * 3 + 1], means[g * 3 + 2]);
        const v3<T> pc = matvec(cam.R, mu) + cam.t;
        const T q[4] = {quats[g * 4], quats[g * 4 + 1], quats[g * 4 + 2], quats[g * 4 + 3]};
        const m33<T> Rq = quat_rot(q[0], q[1], q[2], q[3]);
        const T s0 = scales[g * 3], s1 = scales[g * 3 + 1];
        const T *Tm = rt + idx * 9;
        m33<T> G;  // d/dT (row-major)
#pragma unroll
        for (int k = 0; k < 9; ++k) G.m[k / 3][k % 3] = v_rt[idx * 9 + k];
        G.m[2][2] += v_depths[idx];
        const T vmx = v_means2d[idx * 2], vmy = v_means2d[idx * 2 + 1];
        if (vmx != T(0) || vmy != T(0)) {
            const T f = T(1) / (Tm[6] * Tm[6] + Tm[7] * Tm[7] - Tm[8] * Tm[8]);
            const T sg[3] = {T(1), T(1), T(-1)};
#pragma unroll
            for (int j = 0; j < 3; ++j) {
                G.m[0][j] += vmx * f * sg[j] * Tm[6 + j];
                G.m[1][j] += vmy * f * sg[j] * Tm[6 + j];
            }
            if (mode == TGS_VJP_EXACT) {
                const T mx = f * (Tm[0] * Tm[6] + Tm[1] * Tm[7] - Tm[2] * Tm[8]);
                const T my = f * (Tm[3] * Tm[6] + Tm[4] * Tm[7] - Tm[5] * Tm[8]);
#pragma unroll
                for (int j = 0; j < 3; ++j)
                    G.m[2][j] += sg[j] * f * (vmx * (Tm[j] - T(2) * mx * Tm[6 + j]) + vmy * (Tm[3 + j] - T(2) * my * Tm[6 + j]));
            } else {
                // the reference keeps T0_j T2_j in place of the full numerator (utils.cuh:694-703)
#pragma unroll
                for (int j = 0; j < 3; ++j) {
                    const T t2 = Tm[6 + j];
                    const T corr = j < 2 ? f - T(2) * f * f * t2 * t2 : -(f + T(2) * f * f * t2 * t2);
                    G.m[2][j] += vmx * Tm[j] * corr + vmy * Tm[3 + j] * corr;
                }
            }
        }
        // d/d[a0 a1 pc] = K^T G  (K = [[fx,0,cx],[0,fy,cy],[0,0,1]])
        m33<T> vWH;
#pragma unroll
        for (int j = 0; j < 3; ++j) {
            vWH.m[0][j] = cam.fx * G.m[0][j];
            vWH.m[1][j] = cam.fy * G.m[1][j];
            vWH.m[2][j] = cam.cx * G.m[0][j] + cam.cy * G.m[1][j] + G.m[2][j];
        }
        const v3<T> va0 = mk3(vWH.m[0][0], vWH.m[1][0], vWH.m[2][0]);
        const v3<T> va1 = mk3(vWH.m[0][1], vWH.m[1][1], vWH.m[2][1]);
        const v3<T> vpc = mk3(vWH.m[0][2], vWH.m[1][2], vWH.m[2][2]);
        const m33<T> Rt = transpose(cam.R);
        const v3<T> vRS0 = matvec(Rt, va0), vRS1 = matvec(Rt, va1), vmu = matvec(Rt, vpc);
        const v3<T> nrot = mk3(Rq.m[0][2], Rq.m[1][2], Rq.m[2][2]);
        const v3<T> ncam = matvec(cam.R, nrot);
        const T sgn = dot(ncam, pc) < T(0) ? T(1) : T(-1);
        const v3<T> vn_in = mk3(v_normals[idx * 3], v_normals[idx * 3 + 1], v_normals[idx * 3 + 2]);
        const v3<T> vtn = sgn * matvec(Rt, vn_in);
        m33<T> vRq;
        const T f0 = mode == TGS_VJP_REFERENCE_NORM ? T(1) / s0 : s0;
        const T f1 = mode == TGS_VJP_REFERENCE_NORM ? T(1) / s1 : s1;
        vRq.m[0][0] = vRS0.x * f0; vRq.m[1][0] = vRS0.y * f0; vRq.m[2][0] = vRS0.z * f0;
        vRq.m[0][1] = vRS1.x * f1; vRq.m[1][1] = vRS1.y * f1; vRq.m[2][1] = vRS1.z * f1;
        vRq.m[0][2] = vtn.x; vRq.m[1][2] = vtn.y; vRq.m[2][2] = vtn.z;
        T vq[4];
        quat_rot_vjp(q, vRq, vq);
        const T vs0 = vRS0.x * Rq.m[0][0] + vRS0.y * Rq.m[1][0] + vRS0.z * Rq.m[2][0];
        const T vs1 = vRS1.x * Rq.m[0][1] + vRS1.y * Rq.m[1][1] + vRS1.z * Rq.m[2][1];
        // one camera (or sparse outputs): each Gaussian has at most one entry, so its
        // gradients are stored, not added (the caller zero-fills the outputs)
        if (sparse || C == 1) {
            v_means[og * 3] = vmu.x;
            v_means[og * 3 + 1] = vmu.y;
            v_means[og * 3 + 2] = vmu.z;
#pragma unroll
            for (int k = 0; k < 4; ++k) v_quats[og * 4 + k] = vq[k];
            v_scales[og * 3] = vs0;
            v_scales[og * 3 + 1] = vs1;
        } else {
            atomic_add(v_means + og * 3, vmu.x);
            atomic_add(v_means + og * 3 + 1, vmu.y);
            atomic_add(v_means + og * 3 + 2, vmu.z);
#pragma unroll
            for (int k = 0; k < 4; ++k) atomic_add(v_quats + og * 4 + k, vq[k]);
            atomic_add(v_scales + og * 3, vs0);
            atomic_add(v_scales + og * 3 + 1, vs1);
        }
        if (mode == TGS_VJP_EXACT && v_viewmats != nullptr) {
            // p_c = R mu + t;  [a0 a1] = R Rq diag(s0,s1);  n = sgn R Rq e2
            const v3<T> m0 = mk3(Rq.m[0][0] * s0, Rq.m[1][0] * s0, Rq.m[2][0] * s0);
            const v3<T> m1 = mk3(Rq.m[0][1] * s1, Rq.m[1][1] * s1, Rq.m[2][1] * s1);
            const T va0a[3] = {va0.x, va0.y, va0.z}, va1a[3] = {va1.x, va1.y, va1.z};
            const T vpca[3] = {vpc.x, vpc.y, vpc.z}, vna[3] = {vn_in.x, vn_in.y, vn_in.z};
            const T m0a[3] = {m0.x, m0.y, m0.z}, m1a[3] = {m1.x, m1.y, m1.z};
            const T mua[3] = {mu.x, mu.y, mu.z}, nra[3] = {nrot.x, nrot.y, nrot.z};
#pragma unroll
            for (int r = 0; r < 3; ++r) {
#pragma unroll
                for (int k = 0; k < 3; ++k)
                    vR[r][k] = va0a[r] * m0a[k] + va1a[r] * m1a[k] + vpca[r] * mua[k] + sgn * vna[r] * nra[k];
                vR[r][3] = vpca[r];
            }
        }
    }
    if (mode == TGS_VJP_EXACT && v_viewmats != nullptr) {
        // per-camera reduction: the lanes of one wave may span several cameras (entries
        // are in ascending camera order in both layouts)
        const int64_t base = (int64_t)blockIdx.x * blockDim.x + (threadIdx.x & ~(kWave - 1));
        const int64_t first = base < E ? base : E - 1;
        const int64_t last = base + kWave - 1 < E ? base + kWave - 1 : E - 1;
        const int64_t c_lo = camera_ids ? camera_ids[first] : first / N;
        const int64_t c_hi = camera_ids ? camera_ids[last] : last / N;
        for (int64_t cc = c_lo; cc <= c_hi && cc < C; ++cc) {
            const bool mine = active && c == cc;
#pragma unroll
            for (int r = 0; r < 3; ++r)
#pragma unroll
                for (int k = 0; k < 4; ++k) {
                    const T s = wave_sum(mine ? vR[r][k] : T(0));
                    if ((threadIdx.x & (kWave - 1)) == 0) atomic_add(v_viewmats + cc * 16 + r * 4 + k, s);
                }
        }
    }
}

template <typename T>
static int proj_fwd(int64_t C, int64_t N, const void *means, const void *quats, const void *scales,
                    const void *viewmats, const void *Ks, int32_t W, int32_t H, double nearp, double farp,
                    double clip, int32_t *radii, void *means2d, void *depths, void *rt, void *normals,
                    hipStream_t st) {
    const int64_t n = C * N;
    if (n == 0) return TGS_OK;
    projection_fwd_kernel<T><<<(unsigned)ceil_div(n, 256), 256, 0, st>>>(
        C, N, (const T *)means, (const T *)quats, (const T *)scales, (const T *)viewmats, (const T *)Ks, W, H,
        (T)nearp, (T)farp, (T)clip, radii, (T *)means2d, (T *)depths, (T *)rt, (T *)normals);
    TGS_CHECK_LAUNCH("projection_2dgs_fwd");
    return TGS_OK;
}

template <typename T>
static int proj_bwd(int64_t C, int64_t N, int64_t E, const int64_t *camera_ids, const int64_t *gaussian_ids,
                    int sparse, const void *means, const void *quats, const void *scales,
                    const void *viewmats, const void *Ks, const int32_t *radii, const void *rt,
                    const void *vm2, const void *vd, const void *vn, const void *vrt, int mode, void *v_means,
                    void *v_quats, void *v_scales, void *v_viewmats, hipStream_t st) {
    if (E == 0) return TGS_OK;
    projection_bwd_kernel<T><<<(unsigned)ceil_div(E, 256), 256, 0, st>>>(
        C, N, E, camera_ids, gaussian_ids, sparse, (const T *)means, (const T *)quats, (const T *)scales,
        (const T *)viewmats, (const T *)Ks, radii,
        (const T *)rt, (const T *)vm2, (const T *)vd, (const T *)vn, (const T *)vrt, mode, (T *)v_means,
        (T *)v_quats, (T *)v_scales, (T *)v_viewmats);
    TGS_CHECK_LAUNCH("projection_2dgs_bwd");
    return TGS_OK;
}

template <typename T>
static void proj_flags(int64_t C, int64_t N, const void *means, const void *quats, const void *scales,
                       const void *viewmats, const void *Ks, int32_t W, int32_t H, double nearp, double farp,
                       double clip, int32_t *flags, hipStream_t st) {
    projection_flag_kernel<T><<<(unsigned)ceil_div(C * N, 256), 256, 0, st>>>(
        C, N, (const T *)means, (const T *)quats, (const T *)scales, (const T *)viewmats, (const T *)Ks, W, H,
        (T)nearp, (T)farp, (T)clip, flags);
}

template <typename T>
static void proj_packed(int64_t C, int64_t N, const void *means, const void *quats, const void *scales,
                        const void *viewmats, const void *Ks, int32_t W, int32_t H, double nearp, double farp,
                        double clip, const int64_t *incl, int64_t *indptr, int64_t *camera_ids,
                        int64_t *gaussian_ids, int32_t *radii, void *means2d, void *depths, void *rt, void *normals,
                        hipStream_t st) {
    const int64_t n = std::max<int64_t>(C * N, C + 1);  // one thread per entry and per indptr slot
    projection_packed_kernel<T><<<(unsigned)ceil_div(n, 256), 256, 0, st>>>(
        C, N, (const T *)means, (const T *)quats, (const T *)scales, (const T *)viewmats, (const T *)Ks, W, H,
        (T)nearp, (T)farp, (T)clip, incl, indptr, camera_ids, gaussian_ids, radii, (T *)means2d, (T *)depths,
        (T *)rt, (T *)normals);
}

}  // namespace tgs

// defined in isect.hip
extern "C" int tgs_prefix_sum_i32(int64_t n, const int32_t *counts, int64_t *inclusive, void *temp,
                                  size_t temp_bytes, tgs_stream_t stream);

extern "C" int tgs_projection_2dgs_packed_count(int dtype, int64_t C, int64_t N, const void *means,
                                                const void *quats, const void *scales, const void *viewmats,
                                                const void *Ks, int32_t width, int32_t height, double near_plane,
                                                double far_plane, double radius_clip, int32_t *flags,
                                                int64_t *inclusive, void *temp, size_t temp_bytes,
                                                tgs_stream_t stream) {
    TGS_CHECK_ARG(C >= 0 && N >= 0 && C * N < (int64_t(1) << 31), "projection_packed: bad sizes C=%lld N=%lld",
                  (long long)C, (long long)N);
    TGS_CHECK_ARG(dtype == TGS_F32 || dtype == TGS_F64, "projection_packed: unsupported dtype %d", dtype);
    if (C * N == 0) return TGS_OK;
    hipStream_t st = (hipStream_t)stream;
    if (dtype == TGS_F32)
        tgs::proj_flags<float>(C, N, means, quats, scales, viewmats, Ks, width, height, near_plane, far_plane,
                               radius_clip, flags, st);
    else
        tgs::proj_flags<double>(C, N, means, quats, scales, viewmats, Ks, width, height, near_plane, far_plane,
                                radius_clip, flags, st);
    TGS_CHECK_LAUNCH("projection_packed flags");
    return tgs_prefix_sum_i32(C * N, flags, inclusive, temp, temp_bytes, stream);
}

extern "C" int tgs_projection_2dgs_packed_fill(int dtype, int64_t C, int64_t N, const void *means,
                                               const void *quats, const void *scales, const void *viewmats,
                                               const void *Ks, int32_t width, int32_t height, double near_plane,
                                               double far_plane, double radius_clip, const int64_t *inclusive,
                                               int64_t *indptr, int64_t *camera_ids, int64_t *gaussian_ids,
                                               int32_t *radii, void *means2d, void *depths, void *ray_transforms,
                                               void *normals, tgs_stream_t stream) {
    TGS_CHECK_ARG(C >= 0 && N >= 0 && C * N < (int64_t(1) << 31), "projection_packed: bad sizes");
    TGS_CHECK_ARG(dtype == TGS_F32 || dtype == TGS_F64, "projection_packed: unsupported dtype %d", dtype);
    TGS_CHECK_ARG(C * N == 0 || inclusive, "projection_packed: inclusive prefix sum required");
    hipStream_t st = (hipStream_t)stream;
    if (C * N == 0) return indptr ? tgs_zero_fill(indptr, (size_t)(C + 1) * sizeof(int64_t), stream) : TGS_OK;
    if (dtype == TGS_F32)
        tgs::proj_packed<float>(C, N, means, quats, scales, viewmats, Ks, width, height, near_plane, far_plane,
                                radius_clip, inclusive, indptr, camera_ids, gaussian_ids, radii, means2d, depths,
                                ray_transforms, normals, st);
    else
        tgs::proj_packed<double>(C, N, means, quats, scales, viewmats, Ks, width, height, near_plane, far_plane,
                                 radius_clip, inclusive, indptr, camera_ids, gaussian_ids, radii, means2d, depths,
                                 ray_transforms, normals, st);
    TGS_CHECK_LAUNCH("projection_packed");
    return TGS_OK;
}

extern "C" int tgs_projection_2dgs_packed_bwd(int dtype, int64_t C, int64_t N, int64_t nnz, const int64_t *camera_ids,
                                              const int64_t *gaussian_ids, int sparse, const void *means,
                                              const void *quats, const void *scales, const void *viewmats,
                                              const void *Ks, const int32_t *radii, const void *ray_transforms,
                                              const void *v_means2d, const void *v_depths, const void *v_normals,
                                              const void *v_ray_transforms, int vjp_mode, void *v_means,
                                              void *v_quats, void *v_scales, void *v_viewmats,
                                              tgs_stream_t stream) {
    TGS_CHECK_ARG(C >= 0 && N >= 0 && nnz >= 0 && nnz < (int64_t(1) << 31), "projection_packed_bwd: bad sizes");
    TGS_CHECK_ARG(nnz == 0 || (camera_ids && gaussian_ids), "projection_packed_bwd: ids required");
    TGS_CHECK_ARG(vjp_mode >= 0 && vjp_mode <= 2, "projection_packed_bwd: bad vjp_mode %d", vjp_mode);
    hipStream_t st = (hipStream_t)stream;
    if (dtype == TGS_F32)
        return tgs::proj_bwd<float>(C, N, nnz, camera_ids, gaussian_ids, sparse, means, quats, scales, viewmats, Ks,
                                    radii, ray_transforms, v_means2d, v_depths, v_normals, v_ray_transforms, vjp_mode,
                                    v_means, v_quats, v_scales, v_viewmats, st);
    if (dtype == TGS_F64)
        return tgs::proj_bwd<double>(C, N, nnz, camera_ids, gaussian_ids, sparse, means, quats, scales, viewmats, Ks,
                                     radii, ray_transforms, v_means2d, v_depths, v_normals, v_ray_transforms,
                                     vjp_mode, v_means, v_quats, v_scales, v_viewmats, st);
    tgs::set_error("projection_packed_bwd: unsupported dtype %d", dtype);
    return TGS_ERR_UNSUPPORTED;
}

extern "C" int tgs_projection_2dgs_fwd(int dtype, int64_t C, int64_t N, const void *means, const void *quats,
                                       const void *scales, const void *viewmats, const void *Ks, int32_t width,
                                       int32_t height, double near_plane, double far_plane, double radius_clip,
                                       int32_t *radii, void *means2d, void *depths, void *ray_transforms,
                                       void *normals, tgs_stream_t stream) {
    TGS_CHECK_ARG(C >= 0 && N >= 0 && C * N < (int64_t(1) << 31), "projection_2dgs_fwd: bad sizes C=%lld N=%lld",
                  (long long)C, (long long)N);
    hipStream_t st = (hipStream_t)stream;
    if (dtype == TGS_F32)
        return tgs::proj_fwd<float>(C, N, means, quats, scales, viewmats, Ks, width, height, near_plane, far_plane,
                                    radius_clip, radii, means2d, depths, ray_transforms, normals, st);
    if (dtype == TGS_F64)
        return tgs::proj_fwd<double>(C, N, means, quats, scales, viewmats, Ks, width, height, near_plane, far_plane,
                                     radius_clip, radii, means2d, depths, ray_transforms, normals, st);
    tgs::set_error("projection_2dgs_fwd: unsupported dtype %d", dtype);
    return TGS_ERR_UNSUPPORTED;
}

extern "C" int tgs_projection_2dgs_bwd(int dtype, int64_t C, int64_t N, const void *means, const void *quats,
                                       const void *scales, const void *viewmats, const void *Ks,
                                       const int32_t *radii, const void *ray_transforms, const void *v_means2d,
                                       const void *v_depths, const void *v_normals, const void *v_ray_transforms,
                                       int vjp_mode, void *v_means, void *v_quats, void *v_scales, void *v_viewmats,
                                       tgs_stream_t stream) {
    TGS_CHECK_ARG(C >= 0 && N >= 0 && C * N < (int64_t(1) << 31), "projection_2dgs_bwd: bad sizes");
    TGS_CHECK_ARG(vjp_mode >= 0 && vjp_mode <= 2, "projection_2dgs_bwd: bad vjp_mode %d", vjp_mode);
    hipStream_t st = (hipStream_t)stream;
    if (dtype == TGS_F32)
        return tgs::proj_bwd<float>(C, N, C * N, nullptr, nullptr, 0, means, quats, scales, viewmats, Ks, radii,
                                    ray_transforms, v_means2d, v_depths, v_normals, v_ray_transforms, vjp_mode,
                                    v_means, v_quats, v_scales, v_viewmats, st);
    if (dtype == TGS_F64)
        return tgs::proj_bwd<double>(C, N, C * N, nullptr, nullptr, 0, means, quats, scales, viewmats, Ks, radii,
                                     ray_transforms, v_means2d, v_depths, v_normals, v_ray_transforms, vjp_mode,
                                     v_means, v_quats, v_scales, v_viewmats, st);
    tgs::set_error("projection_2dgs_bwd: unsupported dtype %d", dtype);
    return TGS_ERR_UNSUPPORTED;
}
