// Images of point sources given in the lens frame, and the sequential (warm-started)
// variant that follows the images along a sequence of source positions
// (reference: _images_point_source, src/caustics/point_source.py:1652-1709, and
// _images_point_source_sequential, :1711-1759).
//
// Independent form: one lane per source position (from the caller's initial roots when
// given, the reference's custom_init).  Sequential form: one lane per sequence; the first
// position is solved from the default initial estimates (or the caller's), every later
// one from the previous position's roots, so that root j follows one image along the
// sequence as the reference's lax.scan does.  Images are real when
// |lens_eq(z) - w| < 1e-6 (:1703-1707).
#include "ps_impl.hpp"

using namespace caus;

namespace {

constexpr int kImgBlock = 256;

// lens-frame parameters (a, e1) or (a, e1, e2, Re r3, Im r3), as the reference's private
// image functions take them
template <int NL>
__device__ __forceinline__ void frame_from_lens_params(const double* p, LensFrame<NL>& lf) {
    lf.r[0] = cd{p[0], 0.0};
    lf.r[1] = cd{-p[0], 0.0};
    lf.e[0] = p[1];
    if constexpr (NL == 2) {
        lf.e[1] = 1.0 - p[1];
    } else {
        lf.e[1] = p[2];
        lf.e[2] = 1.0 - p[1] - p[2];
        lf.r[2] = cd{p[3], p[4]};
    }
}

template <int NL, int MODE>
__device__ __forceinline__ bool images_at(cd w, const LensFrame<NL>& lf, cd (&z)[ps_degree<NL>()],
                                          bool warm, int itmax, uint8_t* mask) {
    constexpr int DEG = ps_degree<NL>();
    bool ok = true;
    if constexpr (NL == 1) {
        // closed form (:1664-1671): every image is real
        const double sq = sqrt(1.0 + 4.0 * rcp(norm2(w)));
        z[0] = (0.5 * (1.0 + sq)) * w;
        z[1] = (0.5 * (1.0 - sq)) * w;
        mask[0] = mask[1] = 1;
    } else {
        ok = solve_lens<NL, MODE>(w, lf, z, itmax, warm);
#pragma unroll
        for (int j = 0; j < DEG; ++j) {
            const ImageTerms<NL> it = image_terms<NL>(z[j], lf);
            mask[j] = norm2(z[j] - it.G - w) < 1e-12 ? 1 : 0;
        }
    }
    return ok;
}

template <int NL, int MODE>
__global__ __launch_bounds__(kImgBlock) void k_images(const cd* __restrict__ w, int64_t B,
                                                      int64_t N, const double* lens,
                                                      int64_t lstride, const cd* z_init,
                                                      cd* z_out, uint8_t* mask_out,
                                                      uint8_t* status, int itmax, int sequential) {
    constexpr int DEG = ps_degree<NL>();
    const int64_t t = (int64_t)blockIdx.x * kImgBlock + threadIdx.x;
    if (sequential ? t >= B : t >= B * N) return;
    const int64_t b = sequential ? t : t / N;
    LensFrame<NL> lf;
    if constexpr (NL > 1) frame_from_lens_params<NL>(lens + b * lstride, lf);
    cd z[DEG];
    uint8_t m[DEG];
    const int64_t k0 = sequential ? b * N : t, k1 = sequential ? (b + 1) * N : t + 1;
    for (int64_t i = k0; i < k1; ++i) {
        const bool first = (i == k0);
        const bool warm = !first || z_init != nullptr;
        if (first && z_init) {
#pragma unroll
            for (int j = 0; j < DEG; ++j) z[j] = z_init[i * DEG + j];
        }
        const bool ok = images_at<NL, MODE>(w[i], lf, z, warm, itmax, m);
#pragma unroll
        for (int j = 0; j < DEG; ++j) {
            z_out[i * DEG + j] = z[j];
            mask_out[i * DEG + j] = m[j];
        }
        if (status) status[i] = ok ? 1 : 0;
    }
}

template <int NL>
int images_launch(hipStream_t s, const cd* w, int64_t B, int64_t N, const double* lens,
                  int64_t lstride, const cd* z_init, cd* z, uint8_t* mask, uint8_t* status,
                  int itmax, int mode, int sequential) {
    const int64_t lanes = sequential ? B : B * N;
    const dim3 grid((unsigned)((lanes + kImgBlock - 1) / kImgBlock)), block(kImgBlock);
    switch (mode) {
        case CAUS_MODE_FP64:
            hipLaunchKernelGGL((k_images<NL, MODE_FP64>), grid, block, 0, s, w, B, N, lens, lstride,
                               z_init, z, mask, status, itmax, sequential);
            break;
        case CAUS_MODE_FP64_COMP:
            hipLaunchKernelGGL((k_images<NL, MODE_FP64_COMP>), grid, block, 0, s, w, B, N, lens,
                               lstride, z_init, z, mask, status, itmax, sequential);
            break;
        default:
            hipLaunchKernelGGL((k_images<NL, MODE_CF32>), grid, block, 0, s, w, B, N, lens, lstride,
                               z_init, z, mask, status, itmax, sequential);
            break;
    }
    return hipGetLastError() == hipSuccess ? CAUS_OK : CAUS_ERR_LAUNCH;
}

}  // namespace

extern "C" int caus_images_point_source(void* stream, int32_t nlenses, const double* w, int64_t B,
                                        int64_t N, const double* lens, int64_t lens_stride,
                                        const double* z_init, double* z, uint8_t* mask,
                                        uint8_t* status, int32_t itmax, int32_t mode,
                                        int32_t sequential) {
    if (B < 0 || N < 0 || itmax < 0 || nlenses < 1 || nlenses > 3 || lens_stride < 0 ||
        (sequential != 0 && sequential != 1))
        return CAUS_ERR_ARG;
    if (mode != CAUS_MODE_FP64 && mode != CAUS_MODE_FP64_COMP && mode != CAUS_MODE_CF32)
        return CAUS_ERR_ARG;
    if (B == 0 || N == 0) return CAUS_OK;
    if ((sequential ? B : B * N) > (int64_t(1) << 31) - kImgBlock || !w || !z || !mask ||
        (nlenses > 1 && !lens))
        return CAUS_ERR_ARG;
    hipStream_t s = (hipStream_t)stream;
    const cd* wc = reinterpret_cast<const cd*>(w);
    const cd* zi = reinterpret_cast<const cd*>(z_init);
    cd* zc = reinterpret_cast<cd*>(z);
    switch (nlenses) {
        case 1:
            return images_launch<1>(s, wc, B, N, lens, lens_stride, zi, zc, mask, status, itmax,
                                    mode, sequential);
        case 2:
            return images_launch<2>(s, wc, B, N, lens, lens_stride, zi, zc, mask, status, itmax,
                                    mode, sequential);
        default:
            return images_launch<3>(s, wc, B, N, lens, lens_stride, zi, zc, mask, status, itmax,
                                    mode, sequential);
    }
}
