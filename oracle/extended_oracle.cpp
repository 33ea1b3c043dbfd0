// C++ FP64 restatement of the extended-source contour route, with OpenMP over sources.
// TEST INFRASTRUCTURE ONLY: the host CPU baseline of the extended-source workloads and a
// cross-check of oracle/extended.py, which it follows function by function (file:line in
// the reference, fbartolic/caustics, src/caustics/):
//   limb images, refinement ........ extended_source.py:56-153 (images_of_source_limb)
//   _permute_images, match_points .. extended_source.py:34-53, utils.py:15-40
//   _split_segment ................. extended_source.py:156-216
//   _process_segments, _get_segments extended_source.py:220-309
//   _concatenate_segments, _get_segment_length, _connection_condition, _merge_two_segments,
//   _merge_open_segments ........... extended_source.py:312-667
//   _contours_from_*_segments ...... extended_source.py:670-727
//   mag_extended_source ............ extended_source.py:730-901
//   _integrate_unif, _integrate_ld, _brightness_profile, trapz_zero_avoiding
//                                    integrate.py:18-121, utils.py:43-100
// The roots come from ora_ehrlich_aberth (ea_oracle.c, the reference's initial estimates and
// stopping test) and the lens polynomials from lens_oracle.c's formulas; the Gauss-Legendre
// nodes are passed in (numpy's leggauss).  Sums run in index order (NumPy's pairwise sums
// differ by rounding only).
#include <algorithm>
#include <cmath>
#include <complex>
#include <cstdint>
#include <cstring>
#include <numeric>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

using dc = std::complex<double>;
struct cplx {
    double re, im;
};

extern "C" int ora_ehrlich_aberth(int deg, int itmax, int compensated, int custom_init,
                                  int init_mode, const cplx *poly, const cplx *roots_init,
                                  cplx *roots, int64_t *stats);
extern "C" void ora_polish(int deg, const cplx *p, cplx *z);

// ---- lens ----------------------------------------------------------------------------
struct Lens {
    int nl = 2, n = 5;       // lenses, images (polynomial degree)
    bool polish = false;     // end every limb solve with ora_polish (canonical roots)
    double a = 0, x_cm = 0;
    dc r[3];
    double e[3];
};

Lens make_lens(int nl, const double *prm) {
    Lens L;
    L.nl = nl;
    L.n = nl * nl + 1;
    const double s = prm[0], q = prm[1];
    L.a = 0.5 * s;
    L.x_cm = L.a * (1 - q) / (1 + q);
    if (nl == 2) {
        const double e1 = 1 / (1 + q);
        L.r[0] = L.a;
        L.r[1] = -L.a;
        L.e[0] = e1;
        L.e[1] = 1.0 - e1;
    } else {
        const double q3 = prm[2], r3 = prm[3], psi = prm[4];
        const double e1 = q / (1 + q + q3), e2 = q * e1;
        L.r[0] = L.a;
        L.r[1] = -L.a;
        L.r[2] = r3 * std::exp(dc(0.0, psi));
        L.e[0] = e1;
        L.e[1] = e2;
        L.e[2] = 1.0 - e1 - e2;
    }
    return L;
}

// point_source.py:17-86 (as oracle/lens.py restates it), ascending
void coeffs_binary(dc w, double a, double e1, dc *asc) {
    const dc wb = std::conj(w);
    const double a2 = a * a, a3 = std::pow(a, 3.0), a4 = std::pow(a, 4.0),
                 a5 = std::pow(a, 5.0), a6 = std::pow(a, 6.0);
    const dc wb2 = wb * wb;
    dc p[6];
    p[0] = -a2 + wb2;
    p[1] = a2 * w - 2 * a * e1 + a - w * wb2 + wb;
    p[2] = 2 * a4 - 2 * a2 * wb2 + 4 * a * wb * e1 - 2 * a * wb - 2.0 * w * wb;
    p[3] = (-2 * a4 * w + 4 * a3 * e1 - 2 * a3 + 2 * a2 * w * wb2 - 4 * a * w * wb * e1 +
            2 * a * w * wb + 2 * a * e1 - a - w);
    p[4] = (-a6 + a4 * wb2 - 4 * a3 * wb * e1 + 2 * a3 * wb + 2 * a2 * w * wb +
            4 * a2 * e1 * e1 - 4 * a2 * e1 + 2 * a2 - 4 * a * w * e1 + 2 * a * w);
    p[5] = (a6 * w - 2 * a5 * e1 + a5 - a4 * w * wb2 - a4 * wb + 4 * a3 * w * wb * e1 -
            2 * a3 * w * wb + 2 * a3 * e1 - a3 - 4 * a2 * w * e1 * e1 + 4 * a2 * w * e1 - a2 * w);
    for (int i = 0; i < 6; ++i) asc[i] = p[5 - i];
}

int pmul(const dc *x, int nx, const dc *y, int ny, dc *out) {
    for (int i = 0; i < nx + ny - 1; ++i) out[i] = 0;
    for (int i = 0; i < nx; ++i)
        for (int j = 0; j < ny; ++j) out[i + j] += x[i] * y[j];
    return nx + ny - 1;
}

// oracle/lens.py: poly_coeffs_product for three lenses, ascending (degree 10)
void coeffs_triple(dc w, const Lens &L, dc *P) {
    dc lin[3][2], D[4], t2[3], tmp[7], Nn[3], M[3][4], M01[7], MM[10], zw[2], Pf[11], S[7],
        DS[10], ef[4];
    for (int j = 0; j < 3; ++j) {
        lin[j][0] = -L.r[j];
        lin[j][1] = 1;
    }
    pmul(lin[0], 2, lin[1], 2, t2);
    pmul(t2, 3, lin[2], 2, D);
    for (int i = 0; i < 3; ++i) Nn[i] = 0;
    for (int j = 0; j < 3; ++j) {
        const int k1 = j == 0 ? 1 : 0, k2 = j == 2 ? 1 : 2;
        for (int i = 0; i < 2; ++i) ef[i] = L.e[j] * lin[k1][i];
        pmul(ef, 2, lin[k2], 2, tmp);
        for (int i = 0; i < 3; ++i) Nn[i] += tmp[i];
    }
    const dc wb = std::conj(w);
    for (int j = 0; j < 3; ++j) {
        const dc f = wb - std::conj(L.r[j]);
        for (int i = 0; i < 4; ++i) M[j][i] = f * D[i] + (i < 3 ? Nn[i] : dc(0));
    }
    pmul(M[0], 4, M[1], 4, M01);
    pmul(M01, 7, M[2], 4, MM);
    zw[0] = -w;
    zw[1] = 1;
    pmul(zw, 2, MM, 10, Pf);
    for (int i = 0; i < 7; ++i) S[i] = 0;
    for (int j = 0; j < 3; ++j) {
        const int k1 = j == 0 ? 1 : 0, k2 = j == 2 ? 1 : 2;
        for (int i = 0; i < 4; ++i) ef[i] = L.e[j] * M[k1][i];
        pmul(ef, 4, M[k2], 4, tmp);
        for (int i = 0; i < 7; ++i) S[i] += tmp[i];
    }
    pmul(D, 4, S, 7, DS);
    for (int i = 0; i < 11; ++i) P[i] = Pf[i] - (i < 10 ? DS[i] : dc(0));
}

dc lens_eq(dc z, const Lens &L) {
    const dc zb = std::conj(z);
    dc w = z;
    for (int j = 0; j < L.nl; ++j) w -= L.e[j] / (zb - std::conj(L.r[j]));
    return w;
}

double det_jac(dc z, const Lens &L) {
    const dc zb = std::conj(z);
    dc k = 0;
    for (int j = 0; j < L.nl; ++j) {
        const dc d = zb - std::conj(L.r[j]);
        k += L.e[j] / (d * d);
    }
    const double ak = std::abs(k);
    return 1.0 - ak * ak;
}

// images of one limb point: rows z[i], real-image mask, parity sign(det J)
bool solve_point(dc w, const Lens &L, dc *z, double *mask, double *par) {
    dc asc[11];
    if (L.nl == 2)
        coeffs_binary(w, L.a, L.e[0], asc);
    else
        coeffs_triple(w, L, asc);
    cplx poly[11], roots[10];
    std::memcpy(poly, asc, sizeof(cplx) * (L.n + 1));
    const bool ok = ora_ehrlich_aberth(L.n, 2500, 0, 0, 0, poly, nullptr, roots, nullptr) != 0;
    if (L.polish) ora_polish(L.n, poly, roots);
    for (int i = 0; i < L.n; ++i) {
        z[i] = dc(roots[i].re, roots[i].im);
        mask[i] = std::abs(lens_eq(z[i], L) - w) < 1e-6 ? 1.0 : 0.0;
        const double d = det_jac(z[i], L);
        par[i] = d > 0 ? 1.0 : (d < 0 ? -1.0 : 0.0);
    }
    return ok;
}

// utils.py:15-40: for each a_i in order, the nearest b not yet used (stable ties)
void match_points(const dc *a, const dc *b, int n, int *out) {
    bool used[16] = {false};
    int idx[16];
    double d[16];
    for (int i = 0; i < n; ++i) {
        for (int j = 0; j < n; ++j) {
            idx[j] = j;
            d[j] = std::abs(b[j] - a[i]);
        }
        std::stable_sort(idx, idx + n, [&](int x, int y) { return d[x] < d[y]; });
        for (int t = 0; t < n; ++t)
            if (!used[idx[t]]) {
                used[idx[t]] = true;
                out[i] = idx[t];
                break;
            }
    }
}

// limb images as columns: Z[i * N + k] (image row i, limb point k)
struct Limb {
    int n = 0, N = 0;
    std::vector<dc> z;
    std::vector<double> mask, par, theta;
};

void chain_permute(Limb &lb) {
    const int n = lb.n, N = lb.N;
    std::vector<dc> carry(n), col(n);
    for (int i = 0; i < n; ++i) carry[i] = lb.z[i * N];
    int idx[16];
    dc zc[16];
    double mc[16], pc[16];
    for (int k = 0; k < N; ++k) {
        for (int i = 0; i < n; ++i) col[i] = lb.z[i * N + k];
        match_points(carry.data(), col.data(), n, idx);
        for (int i = 0; i < n; ++i) {
            zc[i] = lb.z[idx[i] * N + k];
            mc[i] = lb.mask[idx[i] * N + k];
            pc[i] = lb.par[idx[i] * N + k];
        }
        for (int i = 0; i < n; ++i) {
            lb.z[i * N + k] = zc[i];
            lb.mask[i * N + k] = mc[i];
            lb.par[i * N + k] = pc[i];
            carry[i] = zc[i];
        }
    }
}

bool images_of_source_limb(dc w0, double rho, int npts, const Lens &L, Limb &lb) {
    const int n = L.n;
    const int npts_init = (int)(0.5 * npts);
    const int m = npts_init - 1;
    std::vector<double> theta(npts_init);
    const double pi = 3.14159265358979323846;
    const double step = (pi - -pi) / m;
    for (int k = 0; k < m; ++k) theta[k] = (double)k * step + -pi;   // numpy linspace
    theta[m] = pi - 1e-8;
    bool ok = true;
    auto solve_cols = [&](const std::vector<double> &th, std::vector<dc> &z,
                          std::vector<double> &mk, std::vector<double> &pr) {
        const int Nc = (int)th.size();
        z.assign((size_t)n * Nc, 0);
        mk.assign((size_t)n * Nc, 0);
        pr.assign((size_t)n * Nc, 0);
        dc zz[16];
        double mm[16], pp[16];
        for (int k = 0; k < Nc; ++k) {
            const dc e = std::exp(dc(0.0, th[k]));
            const dc w = rho * e + w0;
            ok = solve_point(w, L, zz, mm, pp) && ok;
            for (int i = 0; i < n; ++i) {
                z[(size_t)i * Nc + k] = zz[i];
                mk[(size_t)i * Nc + k] = mm[i];
                pr[(size_t)i * Nc + k] = pp[i];
            }
        }
    };
    lb.n = n;
    lb.N = npts_init;
    lb.theta = theta;
    solve_cols(theta, lb.z, lb.mask, lb.par);
    chain_permute(lb);
    const int nadd = npts_init / 10;
    for (int it = 0; it < 10; ++it) {
        const int N = lb.N;
        std::vector<double> dmax(N - 1, 0.0);
        for (int k = 0; k + 1 < N; ++k) {
            double mx = 0.0;
            bool first = true;
            for (int i = 0; i < n; ++i) {
                double dz = std::abs(lb.z[i * N + k + 1] - lb.z[i * N + k]);
                if (!(lb.mask[i * N + k + 1] != 0 || lb.mask[i * N + k] != 0)) dz = 0.0;
                mx = first ? dz : std::max(mx, dz);
                first = false;
            }
            dmax[k] = mx;
        }
        // argsort(dmax, stable)[::-1][:nadd]: descending, ties by larger index first
        std::vector<int> ord(N - 1);
        std::iota(ord.begin(), ord.end(), 0);
        std::stable_sort(ord.begin(), ord.end(), [&](int x, int y) { return dmax[x] < dmax[y]; });
        std::reverse(ord.begin(), ord.end());
        std::vector<int> idcs(ord.begin(), ord.begin() + nadd);
        std::vector<double> th_new(nadd);
        for (int c = 0; c < nadd; ++c) th_new[c] = 0.5 * (lb.theta[idcs[c]] + lb.theta[idcs[c] + 1]);
        std::vector<dc> zn;
        std::vector<double> mn, pn;
        solve_cols(th_new, zn, mn, pn);
        int perm[16];
        dc a[16], b[16];
        for (int c = 0; c < nadd; ++c) {
            const int k = idcs[c];
            for (int i = 0; i < n; ++i) {
                a[i] = lb.z[i * N + k];
                b[i] = zn[(size_t)i * nadd + c];
            }
            match_points(a, b, n, perm);
            dc zc[16];
            double mc[16], pc[16];
            for (int i = 0; i < n; ++i) {
                zc[i] = zn[(size_t)perm[i] * nadd + c];
                mc[i] = mn[(size_t)perm[i] * nadd + c];
                pc[i] = pn[(size_t)perm[i] * nadd + c];
            }
            for (int i = 0; i < n; ++i) {
                zn[(size_t)i * nadd + c] = zc[i];
                mn[(size_t)i * nadd + c] = mc[i];
                pn[(size_t)i * nadd + c] = pc[i];
            }
        }
        // np.insert(x, idcs + 1, new): new point c right after point idcs[c]
        std::vector<int> after(N, -1);   // after[k] = c inserted after point k
        for (int c = 0; c < nadd; ++c) after[idcs[c]] = c;
        const int N2 = N + nadd;
        Limb nb;
        nb.n = n;
        nb.N = N2;
        nb.z.assign((size_t)n * N2, 0);
        nb.mask.assign((size_t)n * N2, 0);
        nb.par.assign((size_t)n * N2, 0);
        nb.theta.assign(N2, 0);
        int d = 0;
        for (int k = 0; k < N; ++k) {
            nb.theta[d] = lb.theta[k];
            for (int i = 0; i < n; ++i) {
                nb.z[(size_t)i * N2 + d] = lb.z[i * N + k];
                nb.mask[(size_t)i * N2 + d] = lb.mask[i * N + k];
                nb.par[(size_t)i * N2 + d] = lb.par[i * N + k];
            }
            ++d;
            if (after[k] >= 0) {
                const int c = after[k];
                nb.theta[d] = th_new[c];
                for (int i = 0; i < n; ++i) {
                    nb.z[(size_t)i * N2 + d] = zn[(size_t)i * nadd + c];
                    nb.mask[(size_t)i * N2 + d] = mn[(size_t)i * nadd + c];
                    nb.par[(size_t)i * N2 + d] = pn[(size_t)i * nadd + c];
                }
                ++d;
            }
        }
        lb = std::move(nb);
    }
    chain_permute(lb);
    return ok;
}

// ---- segments (rows z, parity, mask of one image row, length N) ---------------------
struct Seg {
    std::vector<dc> z;     // length Lp (padded)
    std::vector<double> p, m;
};

int first_nonzero_abs(const std::vector<dc> &x) {
    for (size_t k = 0; k < x.size(); ++k)
        if (std::abs(x[k]) != 0) return (int)k;
    return 0;
}
int first_zero_abs(const std::vector<dc> &x) {
    for (size_t k = 0; k < x.size(); ++k)
        if (std::abs(x[k]) == 0) return (int)k;
    return 0;
}
int last_nonzero_real(const std::vector<dc> &x) {
    for (int k = (int)x.size() - 1; k >= 0; --k)
        if (x[k].real() != 0) return k;
    return 0;
}

// extended_source.py:156-216: up to 10 pieces of one row between changepoints
std::vector<Seg> split_segment(const Seg &s) {
    const int N = (int)s.z.size();
    std::vector<int> starts, ends;
    for (int c = 0; c <= N; ++c) {
        bool cp;
        double dm;
        if (c == 0) {
            cp = s.m[0] != 0;
            dm = s.m[0] == 0 ? 0.0 : 1.0;
        } else if (c == N) {
            cp = s.m[N - 1] != 0;
            dm = s.m[N - 1] == 0 ? 0.0 : -1.0;
        } else {
            const dc zd = s.z[c] - s.z[c - 1];
            const bool ddist = zd.real() * zd.real() + zd.imag() * zd.imag() > 0.1 * 0.1;
            const double dpar = s.p[c] - s.p[c - 1], dmask = s.m[c] - s.m[c - 1];
            cp = ddist || dpar != 0 || dmask != 0;
            dm = dmask;
        }
        if (cp && dm >= 0) starts.push_back(c);
        if (cp && dm <= 0) ends.push_back(c);
    }
    starts.resize(std::max<size_t>(starts.size(), 10), 0);
    ends.resize(std::max<size_t>(ends.size(), 10), 0);
    std::vector<Seg> out(10);
    for (int k = 0; k < 10; ++k) {
        const int l = starts[k], r = ends[k];
        Seg o;
        o.z.assign(N, 0);
        o.p.assign(N, 0);
        o.m.assign(N, 0);
        if (!(l == 0 && r == 0))
            for (int t = std::max(l, 0); t < std::min(r, N); ++t) {
                o.z[t] = s.z[t];
                o.p[t] = s.p[t];
                o.m[t] = s.m[t];
            }
        out[k] = std::move(o);
    }
    return out;
}

// extended_source.py:220-248 (as oracle/extended.py: process_segments)
std::vector<Seg> process_segments(const std::vector<Seg> &rows, int nr) {
    std::vector<Seg> segs;
    for (const Seg &r : rows) {
        auto parts = split_segment(r);
        for (auto &p : parts) segs.push_back(std::move(p));
    }
    const int N = rows.empty() ? 0 : (int)rows[0].z.size();
    std::vector<int> nonempty(segs.size());
    for (size_t i = 0; i < segs.size(); ++i) {
        int cnt = 0;
        for (int t = 0; t < N; ++t) cnt += segs[i].z[t] != dc(0) ? 1 : 0;
        if (cnt < 2) {
            std::fill(segs[i].z.begin(), segs[i].z.end(), dc(0));
            std::fill(segs[i].p.begin(), segs[i].p.end(), 0.0);
            std::fill(segs[i].m.begin(), segs[i].m.end(), 0.0);
        }
        bool ne = false;
        for (int t = 0; t < N; ++t) ne = ne || std::abs(segs[i].z[t]) > 0.0;
        nonempty[i] = ne ? 1 : 0;
    }
    std::vector<int> ord(segs.size());
    std::iota(ord.begin(), ord.end(), 0);
    std::stable_sort(ord.begin(), ord.end(), [&](int x, int y) { return nonempty[x] < nonempty[y]; });
    std::reverse(ord.begin(), ord.end());
    std::vector<Seg> out;
    for (int i = 0; i < nr && i < (int)ord.size(); ++i) {
        Seg s = segs[ord[i]];
        const int h = first_nonzero_abs(s.z);
        std::rotate(s.z.begin(), s.z.begin() + h, s.z.end());
        std::rotate(s.p.begin(), s.p.begin() + h, s.p.end());
        std::rotate(s.m.begin(), s.m.begin() + h, s.m.end());
        out.push_back(std::move(s));
    }
    return out;
}

// ---- stitching (extended_source.py:312-667) -----------------------------------------
double segment_length(const Seg &s, int t) {
    double L = 0.0;
    for (size_t k = 0; k + 1 < s.z.size(); ++k)
        if ((int)k != t) L += std::abs(s.z[k + 1] - s.z[k]);
    return L;
}

void head_line(const Seg &s, int t, dc &l0, dc &l1) {
    const std::vector<dc> &x = s.z;
    const bool cond = std::abs(x[1] - x[0]) > 1e-5 || t <= 1;
    if (cond) {
        l0 = x[1];
        l1 = x[0];
    } else {
        l0 = x[2];
        l1 = x[1];
    }
}

void slice2(const std::vector<dc> &x, int start, dc &l0, dc &l1) {
    const int s = std::min(std::max(start, 0), (int)x.size() - 2);
    l0 = x[s];
    l1 = x[s + 1];
}

void tail_line(const Seg &s, int t, dc &l0, dc &l1) {
    const std::vector<dc> &x = s.z;
    const int n = (int)x.size();
    const dc xm1 = x[((t - 1) % n + n) % n];   // numpy negative index wraps
    const bool cond = std::abs(x[t] - xm1) > 1e-5 || t <= 1;
    slice2(x, cond ? t - 1 : t - 2, l0, l1);
}

bool connection_condition(const Seg &s1, const Seg &s2, int t1, int t2, int ctype) {
    const bool same = s1.p[0] * s2.p[0] > 0.0;
    const bool cpar = ctype < 2 ? same : !same;
    dc a0, a1, b0, b1;
    if (ctype == 0) {
        tail_line(s1, t1, a0, a1);
        head_line(s2, t2, b0, b1);
    } else if (ctype == 1) {
        head_line(s1, t1, a0, a1);
        tail_line(s2, t2, b0, b1);
    } else if (ctype == 2) {
        head_line(s1, t1, a0, a1);
        head_line(s2, t2, b0, b1);
    } else {
        tail_line(s1, t1, a0, a1);
        tail_line(s2, t2, b0, b1);
    }
    const double dist = std::abs(a1 - b1);
    const bool c1 = dist < 1e-1;
    const dc v1 = (a1 - a0) / std::abs(a1 - a0), v2 = (b1 - b0) / std::abs(b1 - b0);
    const double alpha = std::acos(v1.real() * v2.real() + v1.imag() * v2.imag());
    const bool c2 = (180.0 - alpha * (180.0 / 3.14159265358979323846)) < 60.0;
    const bool c3 = std::abs(a1 - b1) < std::abs(a0 - b0);
    return cpar && ((c1 && c2 && c3) || dist < 1e-5);
}

Seg concatenate_segments(const Seg &a, const Seg &b) {
    const int L = first_zero_abs(a.z);
    const int n = (int)a.z.size();
    Seg m = a;
    for (int k = 0; k < n; ++k) {
        const int src = ((k - L) % n + n) % n;   // np.roll(b, L)
        m.z[k] += b.z[src];
        m.p[k] += b.p[src];
    }
    return m;
}

Seg merge_two(const Seg &s1, Seg s2, int t1, int t2, int ctype, int &tout) {
    const int n = (int)s2.z.size();
    if (ctype == 2 || ctype == 3) {
        std::reverse(s2.z.begin(), s2.z.end());
        std::reverse(s2.p.begin(), s2.p.end());
        const int sh = n - t2 - 1;   // np.roll(s2, -sh)
        std::rotate(s2.z.begin(), s2.z.begin() + ((sh % n) + n) % n, s2.z.end());
        std::rotate(s2.p.begin(), s2.p.begin() + ((sh % n) + n) % n, s2.p.end());
        for (double &v : s2.p) v = -v;
    }
    tout = t1 + t2 + 1;
    return (ctype == 0 || ctype == 3) ? concatenate_segments(s1, s2) : concatenate_segments(s2, s1);
}

std::vector<Seg> merge_open_segments(std::vector<Seg> segs) {
    const int N = (int)segs[0].z.size();
    std::vector<int> tails(segs.size());
    for (size_t i = 0; i < segs.size(); ++i) {
        tails[i] = last_nonzero_real(segs[i].z);
        segs[i].z.resize(4 * N, 0);
        segs[i].p.resize(4 * N, 0);
        segs[i].m.clear();
    }
    std::vector<Seg> merged;
    int maxm = 20;
    for (int ci = 0; ci < 3; ++ci) {
        const int R = (int)segs.size();
        std::vector<double> key(R);
        for (int i = 0; i < R; ++i) {
            const double l = segment_length(segs[i], tails[i]);
            key[i] = l != 0 ? l : NAN;
        }
        std::vector<int> ord(R);
        std::iota(ord.begin(), ord.end(), 0);
        std::stable_sort(ord.begin(), ord.end(), [&](int x, int y) {   // NaN last
            const bool nx = std::isnan(key[x]), ny = std::isnan(key[y]);
            if (nx != ny) return ny;
            if (nx) return false;
            return key[x] < key[y];
        });
        std::vector<Seg> s2;
        std::vector<int> t2;
        for (int i : ord) {
            s2.push_back(std::move(segs[i]));
            t2.push_back(tails[i]);
        }
        Seg act = s2[0];
        int tact = t2[0];
        std::vector<Seg> rest(s2.begin() + 1, s2.end());
        std::vector<int> trest(t2.begin() + 1, t2.end());
        const int Rr = (int)rest.size();
        for (int it = 0; it < maxm; ++it) {
            bool any = false;
            for (int i = 0; i < Rr; ++i) any = any || rest[i].z[0] != dc(0);
            if (!any) continue;
            std::vector<double> dist(4 * Rr);
            for (int i = 0; i < Rr; ++i) {
                dist[0 * Rr + i] = std::abs(act.z[tact] - rest[i].z[0]);
                dist[1 * Rr + i] = std::abs(act.z[0] - rest[i].z[trest[i]]);
                dist[2 * Rr + i] = std::abs(act.z[0] - rest[i].z[0]);
                dist[3 * Rr + i] = std::abs(act.z[tact] - rest[i].z[trest[i]]);
            }
            std::vector<int> fl(4 * Rr);
            std::iota(fl.begin(), fl.end(), 0);
            std::stable_sort(fl.begin(), fl.end(), [&](int x, int y) { return dist[x] < dist[y]; });
            for (int u = 0; u < 4 && u < (int)fl.size(); ++u) {
                const int ct = fl[u] / Rr, i = fl[u] % Rr;
                if (connection_condition(act, rest[i], tact, trest[i], ct)) {
                    int tn;
                    act = merge_two(act, rest[i], tact, trest[i], ct, tn);
                    tact = tn;
                    std::fill(rest[i].z.begin(), rest[i].z.end(), dc(0));
                    std::fill(rest[i].p.begin(), rest[i].p.end(), 0.0);
                    break;
                }
            }
        }
        merged.push_back(act);
        segs = std::move(rest);
        tails = std::move(trest);
        maxm -= 2;
    }
    return merged;
}

// ---- integrals (integrate.py:18-121) ---------------------------------------------------
double trapz_zero_avoiding(const std::vector<double> &y, const std::vector<double> &x, int tidx) {
    const int n = (int)x.size();
    double I = 0.0;
    for (int k = 0; k + 1 < n; ++k) I += 0.5 * (y[k + 1] + y[k]) * (x[k + 1] - x[k]);
    if (tidx == n - 1) return I;
    return I - 0.5 * ((y[tidx] + y[tidx + 1]) * (x[tidx + 1] - x[tidx]));
}

double integrate_unif(const std::vector<dc> &z, int tidx) {
    const int n = (int)z.size();
    std::vector<double> a(n), b(n), x(n), y(n);
    for (int k = 0; k < n; ++k) {
        x[k] = z[k].real();
        y[k] = z[k].imag();
        a[k] = 0.5 * x[k];
        b[k] = -0.5 * y[k];
    }
    return trapz_zero_avoiding(a, y, tidx) + trapz_zero_avoiding(b, x, tidx);
}

struct LD {
    double rho, u1;
    dc w0;
    const Lens *L;
    const double *x1, *w1, *x2, *w2;
    int n1, n2;
};

double brightness(dc z, const LD &d) {
    const dc w = lens_eq(z, *d.L);
    const double r = std::abs(w - d.w0) / d.rho;
    double Br;
    if (r <= 1.0) {
        const double v = 1 - r * r;
        Br = 1 + std::sqrt(v > 0 ? v : 0.0);
    } else {
        const double v = 1 - 1.0 / (r * r);
        Br = 1 - std::sqrt(v > 0 ? v : 0.0);
    }
    return 3.0 / (3.0 - d.u1) * (d.u1 * Br + 1.0 - 2.0 * d.u1);
}

double split_point(double a, double b, double rho) {
    const double dd = std::abs(b - a);
    const double sp = b > a ? b - 2 * rho : b + 2 * rho;
    return 0.5 * dd <= 2 * rho ? a + 0.5 * dd : sp;
}

template <class F>
double gl(F f, double a, double b, const double *x, const double *w, int n) {
    double s = 0.0;
    for (int i = 0; i < n; ++i) {
        const double pt = 0.5 * (b - a) * x[i] + 0.5 * (b + a);
        s += 0.5 * (b - a) * f(pt) * w[i];
    }
    return s;
}

double integrate_ld(const std::vector<dc> &z, int tidx, const LD &d) {
    const int n = (int)z.size();
    dc zs = 0;
    for (int k = 0; k < n; ++k) zs += z[k];
    const dc z0 = zs / (double)(tidx + 1);
    std::vector<double> x(n), y(n), P(n), Q(n);
    for (int k = 0; k < n; ++k) {
        x[k] = z[k].real();
        y[k] = z[k].imag();
    }
    for (int k = 0; k < n; ++k) {
        // only contour points and the segment into the padding enter the trapezoid sums
        if (k > tidx + 1) {
            P[k] = Q[k] = 0.0;
            continue;
        }
        {
            const double a = z0.imag(), sp = split_point(a, y[k], d.rho);
            auto f = [&](double yy) { return brightness(dc(x[k], yy), d); };
            P[k] = -0.5 * (gl(f, a, sp, d.x1, d.w1, d.n1) + gl(f, sp, y[k], d.x2, d.w2, d.n2));
        }
        {
            const double a = z0.real(), sp = split_point(a, x[k], d.rho);
            auto f = [&](double xx) { return brightness(dc(xx, y[k]), d); };
            Q[k] = 0.5 * (gl(f, a, sp, d.x1, d.w1, d.n1) + gl(f, sp, x[k], d.x2, d.w2, d.n2));
        }
    }
    return trapz_zero_avoiding(P, x, tidx) + trapz_zero_avoiding(Q, y, tidx);
}

double mag_one(dc w0u, double rho, int nl, const double *prm, int npts, bool ld, double u1,
               const double *x1, const double *w1, int n1, const double *x2, const double *w2,
               int n2, bool polish, bool &ok) {
    Lens L = make_lens(nl, prm);
    L.polish = polish;
    const dc w0 = w0u + L.x_cm;
    Limb lb;
    ok = images_of_source_limb(w0, rho, npts, L, lb);
    const int n = lb.n, N = lb.N;
    LD d{rho, u1, w0, &L, x1, w1, x2, w2, n1, n2};
    auto integ = [&](const std::vector<dc> &c, int t) {
        return ld ? integrate_ld(c, t, d) : integrate_unif(c, t);
    };
    // _get_segments: masked rows, closed rows
    std::vector<Seg> rows(n);
    std::vector<bool> closed(n);
    bool all_closed = true;
    for (int i = 0; i < n; ++i) {
        Seg &s = rows[i];
        s.z.resize(N);
        s.p.resize(N);
        s.m.resize(N);
        bool allm = true;
        for (int k = 0; k < N; ++k) {
            const double m = lb.mask[(size_t)i * N + k];
            s.z[k] = lb.z[(size_t)i * N + k] * m;
            s.p[k] = lb.par[(size_t)i * N + k] * m;
            s.m[k] = m;
            allm = allm && m != 0;
        }
        closed[i] = std::abs(s.z[0] - s.z[N - 1]) < 1e-5 && allm;
        all_closed = all_closed && closed[i];
    }
    double I = 0.0;
    for (int i = 0; i < n; ++i) {
        if (!closed[i]) continue;   // open rows give zero closed contours
        std::vector<dc> c(rows[i].z);
        c.push_back(c[0]);
        I += integ(c, N) * rows[i].p[0];
    }
    if (!all_closed) {
        std::vector<Seg> open(n);
        for (int i = 0; i < n; ++i) {
            open[i] = rows[i];
            if (closed[i]) {
                std::fill(open[i].z.begin(), open[i].z.end(), dc(0));
                std::fill(open[i].p.begin(), open[i].p.end(), 0.0);
                std::fill(open[i].m.begin(), open[i].m.end(), 0.0);
            }
        }
        auto segs = process_segments(open, 3 * n);
        auto m = merge_open_segments(std::move(segs));
        for (int k = 0; k < 3; ++k) {
            std::vector<dc> c = m[k].z;
            const double cp = m[k].p[0];
            const int t = last_nonzero_real(c);
            if (t + 1 < (int)c.size()) c[t + 1] = c[0];
            I += integ(c, last_nonzero_real(c)) * cp;
        }
    }
    return std::abs(I) / (3.14159265358979323846 * rho * rho);
}

}  // namespace

// mag[i] for sources (w0 interleaved complex in the reference's frame, rho, parameter rows
// (s, q[, q3, r3, psi]) at params + i * pstride, npts_limb per source, u1 per source);
// limb darkening with the given Gauss-Legendre nodes of the two half-lines.  polish != 0
// ends every limb solve with the canonical compensated polish (ora_polish; no reference
// counterpart), as oracle/extended.py does with polish=True.  nthreads <= 0: the OpenMP
// default.  Returns the number of sources with an unconverged solve.
extern "C" int64_t ora_mag_extended_source(int nlenses, int64_t S, const double *w0,
                                           const double *rho, const double *params,
                                           int64_t pstride, const int32_t *npts, int ld,
                                           const double *u1, const double *x1, const double *w1,
                                           int n1, const double *x2, const double *w2, int n2,
                                           int polish, double *mag, int nthreads) {
    if (nlenses != 2 && nlenses != 3) return -1;
    int64_t bad = 0;
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel for schedule(dynamic, 1) reduction(+ : bad)
#endif
    for (int64_t i = 0; i < S; ++i) {
        bool ok = true;
        mag[i] = mag_one(dc(w0[2 * i], w0[2 * i + 1]), rho[i], nlenses, params + i * pstride,
                         npts[i], ld != 0, ld ? u1[i] : 0.0, x1, w1, n1, x2, w2, n2,
                         polish != 0, ok);
        bad += ok ? 0 : 1;
    }
    return bad;
}
