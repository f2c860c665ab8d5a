// Batched quasar-continuum fit on gfx950 (MI355X).
//
// Minimises, for every forest independently, the objective of picca's continuum fit
// (LeastsSquaresContModel.__call__,
// py/picca/delta_extraction/least_squares/least_squares_cont_model.py:65-95):
//     chi2(a, b) = sum_{ivar>0} (f - C)^2 W - sum_{ivar>0} log W,   C = (a + b x) m,
// with x = (log_lambda - Lmin) / (Lmax - Lmin), m the mean continuum and
// W = 1 / C^2 / (eta var_pipe + var_lss + fudge / var_pipe), var_pipe = 1 / ivar / C^2
// (:141-192).  Substituting var_pipe, W = 1 / D(C) with
//     D(C) = eta / ivar + var_lss C^2 + fudge ivar C^4,
// so chi2 = sum (f - C)^2 / D + log D: smooth in (a, b), with closed-form first and second
// derivatives.  With use_constant_weight W = 1 and chi2 is the plain sum of squares.
//
// One wave64 per forest: lanes stride over the pixels, the six sums (chi2, gradient,
// Hessian) are reduced across the wave, and every lane runs the same Levenberg-Marquardt
// iteration on them (damped Newton; a step is taken only when it lowers chi2).  The fit
// stops when the Newton decrement g^T H^-1 g / 2 (Migrad's EDM) falls below tol * max(1,
// |chi2|).  The slope stays at 0 for order-0 forests (py/picca/delta_extraction/
// expected_fluxes/utils.py:109).  The validity checks of compute_continuum (:111-120:
// not converged, then negative, then all-zero continuum, each overriding the previous)
// run on the final parameters.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "device_util.h"
#include "engine.h"

namespace picca {

namespace {

struct Sums {
    double chi, ga, gb, haa, hab, hbb;
};

__device__ __forceinline__ double wave_sum(double v) {
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

__device__ Sums evaluate(const ContIn &I, int64_t o, int n, double a, double b, bool constant,
                         int lane) {
    Sums s{0, 0, 0, 0, 0, 0};
    for (int k = lane; k < n; k += 64) {
        const double iv = I.ivar[o + k];
        if (!(iv > 0)) continue;
        const double m = I.mean_cont[o + k];
        const double x = I.x[o + k];
        const double f = I.flux[o + k];
        const double C = (a + b * x) * m;
        const double r = f - C;
        double T, dT, d2T;
        if (constant) {
            T = r * r;
            dT = -2.0 * r;
            d2T = 2.0;
        } else {
            const double eta = I.eta[o + k];
            const double lss = I.var_lss[o + k];
            const double fud = I.fudge[o + k];
            const double C2 = C * C;
            const double D = eta / iv + lss * C2 + fud * iv * C2 * C2;
            const double D1 = 2.0 * lss * C + 4.0 * fud * iv * C2 * C;   // dD/dC
            const double D2 = 2.0 * lss + 12.0 * fud * iv * C2;           // d2D/dC2
            const double inv = 1.0 / D;
            const double q = r * inv;                                     // r / D
            T = r * q + log(D);
            // dT/dC = -2 r / D - r^2 D' / D^2 + D' / D
            dT = -2.0 * q - q * q * D1 + D1 * inv;
            // d2T/dC2 = 2/D + 4 r D'/D^2 + 2 r^2 D'^2/D^3 - r^2 D''/D^2 + D''/D - D'^2/D^2
            d2T = 2.0 * inv + 4.0 * q * D1 * inv + 2.0 * q * q * D1 * D1 * inv -
                  q * q * D2 + D2 * inv - D1 * D1 * inv * inv;
        }
        const double mx = m * x;
        s.chi += T;
        s.ga += dT * m;
        s.gb += dT * mx;
        s.haa += d2T * m * m;
        s.hab += d2T * m * mx;
        s.hbb += d2T * mx * mx;
    }
    s.chi = wave_sum(s.chi);
    s.ga = wave_sum(s.ga);
    s.gb = wave_sum(s.gb);
    s.haa = wave_sum(s.haa);
    s.hab = wave_sum(s.hab);
    s.hbb = wave_sum(s.hbb);
    return s;
}

// Newton (lambda = 0) or damped step; false when the damped Hessian is not positive
__device__ __forceinline__ bool solve_step(const Sums &s, bool free_b, double lambda,
                                           double &da, double &db) {
    const double haa = s.haa * (1.0 + lambda);
    if (!free_b) {
        if (!(haa > 0)) return false;
        da = -s.ga / haa;
        db = 0.0;
        return true;
    }
    const double hbb = s.hbb * (1.0 + lambda);
    const double det = haa * hbb - s.hab * s.hab;
    if (!(haa > 0) || !(det > 0)) return false;
    da = -(hbb * s.ga - s.hab * s.gb) / det;
    db = -(haa * s.gb - s.hab * s.ga) / det;
    return true;
}

}  // namespace

__global__ __launch_bounds__(256) void cont_fit_kernel(ContConst K, ContIn I, ContOut O,
                                                       int64_t n_forests) {
    const int lane = threadIdx.x & 63;
    const int64_t f = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
    if (f >= n_forests) return;   // whole waves leave together
    const int64_t o = I.off[f];
    const int n = (int)(I.off[f + 1] - o);
    const bool free_b = I.order[f] != 0;
    const bool constant = K.use_constant_weight != 0;
    double a = I.a0[f];
    double b = free_b ? I.b0[f] : 0.0;
    Sums s = evaluate(I, o, n, a, b, constant, lane);
    int evals = 1;
    int status = 1;   // 1: not converged within max_iter
    double lambda = 0.0;
    for (int it = 0; it < K.max_iter; ++it) {
        // convergence: Newton decrement at the current point
        double na, nb;
        if (solve_step(s, free_b, 0.0, na, nb)) {
            const double edm = -0.5 * (s.ga * na + s.gb * nb);
            if (edm < K.tol * fmax(1.0, fabs(s.chi)) || edm < K.tol_abs) {
                status = 0;
                break;
            }
        }
        double da, db;
        if (!solve_step(s, free_b, lambda, da, db)) {
            lambda = lambda == 0.0 ? 1e-3 : lambda * 10.0;
            if (lambda > 1e12) break;
            continue;
        }
        const Sums t = evaluate(I, o, n, a + da, b + db, constant, lane);
        ++evals;
        if (t.chi < s.chi) {
            a += da;
            b += db;
            s = t;
            lambda = lambda * 0.1;
            if (lambda < 1e-9) lambda = 0.0;
        } else {
            lambda = lambda == 0.0 ? 1e-3 : lambda * 10.0;
            if (lambda > 1e12) {   // no representable improvement left: at the minimum
                status = 0;
                break;
            }
        }
    }
    // validity of the final continuum over every pixel of the forest
    bool negative = false, nonzero = false;
    int ndata = 0;
    for (int k = lane; k < n; k += 64) {
        const double C = (a + b * I.x[o + k]) * I.mean_cont[o + k];
        negative |= C < 0;
        nonzero |= C != 0;
        ndata += I.ivar[o + k] > 0 ? 1 : 0;
    }
    negative = __ballot(negative) != 0;
    nonzero = __ballot(nonzero) != 0;
    for (int off = 32; off > 0; off >>= 1) ndata += __shfl_xor(ndata, off, 64);
    // the reference's checks in its order, later ones overriding earlier ones
    if (status == 0 && !(s.chi == s.chi)) status = 4;
    if (negative) status = 2;
    if (!nonzero) status = 3;
    if (lane == 0) {
        O.a[f] = a;
        O.b[f] = b;
        O.chi2[f] = s.chi;
        O.ndata[f] = ndata;
        O.status[f] = status;
        O.evals[f] = evals;
    }
}

hipError_t launch_cont_fit(const ContConst &K, const ContIn &I, const ContOut &O,
                           int64_t n_forests, hipStream_t stream) {
    if (n_forests <= 0) return hipSuccess;
    const int64_t blocks = (n_forests + 3) / 4;
    hipLaunchKernelGGL(cont_fit_kernel, dim3((unsigned)blocks), dim3(256), 0, stream, K, I, O,
                       n_forests);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------------------
// Mean-continuum update of the DR16 iteration
// (Dr16ExpectedFlux.compute_expected_flux, py/picca/delta_extraction/expected_fluxes/
// dr16_expected_flux.py:380-464).
//
// cont_stack_kernel: ExpectedFlux._compute_mean_cont (py/picca/delta_extraction/
// expected_flux.py:256-308) with the forest weights of Dr16ExpectedFlux.compute_forest_weights
// (dr16_expected_flux.py:466-488, var_lss_mod = 1): over the pixels of the forests whose
// last fit is valid, sum F / (C + 1e-16) * w and w in the pixel's rest-frame bin, with
// w = 1 / (eta var_pipe + var_lss + fudge / var_pipe), var_pipe = 1 / ivar / C^2 and
// w = 0 where ivar <= 0.  One wave per forest; the bins live in LDS per workgroup and are
// added to HBM once per workgroup.
__global__ void __launch_bounds__(256)
cont_stack_kernel(ContIn I, const double *__restrict__ a_fit, const double *__restrict__ b_fit,
                  const int32_t *__restrict__ status, const int32_t *__restrict__ bins,
                  int32_t nbins, double *__restrict__ sum_ratio,
                  double *__restrict__ sum_weight, int64_t n_forests) {
    extern __shared__ __align__(16) unsigned char smem[];
    double *h_ratio = reinterpret_cast<double *>(smem);
    double *h_weight = h_ratio + nbins;
    for (int k = threadIdx.x; k < nbins; k += blockDim.x) {
        h_ratio[k] = 0.0;
        h_weight[k] = 0.0;
    }
    __syncthreads();
    const int lane = threadIdx.x & 63;
    const int64_t waves = (int64_t)gridDim.x * (blockDim.x >> 6);
    for (int64_t f = (int64_t)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6); f < n_forests;
         f += waves) {
        if (status[f] != 0) continue;
        const double a = a_fit[f], b = b_fit[f];
        const int64_t o = I.off[f];
        const int n = (int)(I.off[f + 1] - o);
        for (int k = lane; k < n; k += 64) {
            const double iv = I.ivar[o + k];
            if (!(iv > 0)) continue;
            const double C = (b * I.x[o + k] + a) * I.mean_cont[o + k];
            const double var_pipe = 1.0 / iv / (C * C);
            const double w = 1.0 / (I.eta[o + k] * var_pipe + I.var_lss[o + k] +
                                    I.fudge[o + k] / var_pipe);
            const double ratio = I.flux[o + k] / (C + 1e-16);
            const int bin = bins[o + k];
            atomicAdd(&h_ratio[bin], ratio * w);
            atomicAdd(&h_weight[bin], w);
        }
    }
    __syncthreads();
    for (int k = threadIdx.x; k < nbins; k += blockDim.x) {
        if (h_weight[k] != 0.0) {
            unsafeAtomicAdd(&sum_ratio[k], h_ratio[k]);
            unsafeAtomicAdd(&sum_weight[k], h_weight[k]);
        }
    }
}

// cont_interp_kernel: the new get_mean_cont, scipy's interp1d(x, y,
// fill_value="extrapolate") (linear, nearest end interval outside [x0, x_{n-1}]), at every
// pixel's rest-frame log lambda, times the transmission correction
// (py/picca/delta_extraction/expected_fluxes/utils.py:49-64).  The interval and the slope
// follow scipy's _call_linear: i = searchsorted(x, v) clipped to [1, n - 1],
// y = (y_i - y_{i-1}) / (x_i - x_{i-1}) * (v - x_{i-1}) + y_{i-1}.
__global__ void __launch_bounds__(256)
cont_interp_kernel(const double *__restrict__ ll_rest, const double *__restrict__ trans,
                   const double *__restrict__ xs, const double *__restrict__ ys, int32_t n,
                   double *__restrict__ mean_cont, int64_t npix) {
    extern __shared__ __align__(16) unsigned char smem[];
    double *sx = reinterpret_cast<double *>(smem);
    double *sy = sx + n;
    for (int k = threadIdx.x; k < n; k += blockDim.x) {
        sx[k] = xs[k];
        sy[k] = ys[k];
    }
    __syncthreads();
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; p < npix; p += stride) {
        const double v = ll_rest[p];
        // first index with sx[i] >= v (numpy searchsorted, side "left")
        int lo = 0, hi = n;
        while (lo < hi) {
            const int mid = (lo + hi) >> 1;
            if (sx[mid] < v) lo = mid + 1;
            else hi = mid;
        }
        const int i = lo < 1 ? 1 : (lo > n - 1 ? n - 1 : lo);
        const double x_lo = sx[i - 1], x_hi = sx[i];
        const double y_lo = sy[i - 1], y_hi = sy[i];
        const double slope = (y_hi - y_lo) / (x_hi - x_lo);
        mean_cont[p] = (slope * (v - x_lo) + y_lo) * trans[p];
    }
}

hipError_t launch_cont_stack(const ContIn &I, const double *a, const double *b,
                             const int32_t *status, const int32_t *bins, int32_t nbins,
                             double *sum_ratio, double *sum_weight, int64_t n_forests,
                             int num_cus, hipStream_t stream) {
    if (n_forests <= 0 || nbins <= 0) return hipSuccess;
    const size_t lds = 2 * sizeof(double) * (size_t)nbins;
    int64_t blocks = (n_forests + 3) / 4;
    if (blocks > 8 * (int64_t)num_cus) blocks = 8 * (int64_t)num_cus;
    hipError_t err = hipFuncSetAttribute((const void *)cont_stack_kernel,
                                         hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if (err != hipSuccess) return err;
    hipLaunchKernelGGL(cont_stack_kernel, dim3((unsigned)blocks), dim3(256), lds, stream, I, a,
                       b, status, bins, nbins, sum_ratio, sum_weight, n_forests);
    return hipGetLastError();
}

hipError_t launch_cont_interp(const double *ll_rest, const double *trans, const double *xs,
                              const double *ys, int32_t n, double *mean_cont, int64_t npix,
                              int num_cus, hipStream_t stream) {
    if (npix <= 0) return hipSuccess;
    const size_t lds = 2 * sizeof(double) * (size_t)n;
    int64_t blocks = (npix + 255) / 256;
    if (blocks > 16 * (int64_t)num_cus) blocks = 16 * (int64_t)num_cus;
    hipError_t err = hipFuncSetAttribute((const void *)cont_interp_kernel,
                                         hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if (err != hipSuccess) return err;
    hipLaunchKernelGGL(cont_interp_kernel, dim3((unsigned)blocks), dim3(256), lds, stream,
                       ll_rest, trans, xs, ys, n, mean_cont, npix);
    return hipGetLastError();
}

}  // namespace picca
