// C ABI of the LBM library: handle lifecycle, device memory, ports, MRT basis, the
// forward sweep and the revolve-driven discrete adjoint (include/implexity_lbm.h).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <memory>
#include <mutex>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/implexity_lbm.h"
#include <hipcub/hipcub.hpp>

#include "design_map.h"
#include "pushforward.h"
#include "jvp.h"
#include "launch.h"
#include "revolve.h"

using namespace lbm;

namespace {

thread_local std::string g_last_error;

struct Fail {
    int code;
    std::string msg;
};

[[noreturn]] void fail(int code, const std::string& msg) { throw Fail{code, msg}; }

void hip_check(hipError_t e, const char* what) {
    if (e != hipSuccess) fail(LBM_ERR_DEVICE, std::string(what) + ": " + hipGetErrorString(e));
}

template <class F>
int guarded(F&& f) {
    try {
        f();
        return LBM_OK;
    } catch (const Fail& e) {
        g_last_error = e.msg;
        return e.code;
    } catch (const std::bad_alloc&) {
        g_last_error = "host allocation failed";
        return LBM_ERR_MEMORY;
    } catch (const std::exception& e) {
        g_last_error = e.what();
        return LBM_ERR_INVALID;
    } catch (...) {
        g_last_error = "unknown failure";
        return LBM_ERR_INVALID;
    }
}

struct DevBuf {
    void* p = nullptr;
    size_t bytes = 0;
    DevBuf() = default;
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
    ~DevBuf() { reset(); }
    void reset() {
        if (p) (void)hipFree(p);
        p = nullptr;
        bytes = 0;
    }
    void alloc(size_t n, const char* what) {
        if (p && bytes == n) return;
        reset();
        if (n == 0) return;
        const hipError_t e = hipMalloc(&p, n);
        if (e != hipSuccess) {
            (void)hipGetLastError();
            p = nullptr;
            fail(LBM_ERR_MEMORY, std::string("device allocation of ") + std::to_string(n) + " bytes for " + what +
                                     " failed: " + hipGetErrorString(e));
        }
        bytes = n;
    }
    // Keeps the allocation when it already holds n bytes (contents are not preserved
    // when it grows).
    void grow(size_t n, const char* what) {
        if (p && bytes >= n) return;
        alloc(n, what);
    }
    template <class T>
    T* as() const {
        return static_cast<T*>(p);
    }
};

// Two pinned host chunks through which large device-to-host copies are staged: chunk
// k + 1 is copied over PCIe while chunk k is copied out of pinned memory on the host.
struct HostStage {
    static constexpr size_t CHUNK = size_t(8) << 20;
    void* buf[2] = {nullptr, nullptr};
    hipEvent_t ev[2] = {nullptr, nullptr};
    HostStage() = default;
    HostStage(const HostStage&) = delete;
    HostStage& operator=(const HostStage&) = delete;
    ~HostStage() {
        for (int b = 0; b < 2; b++) {
            if (buf[b]) (void)hipHostFree(buf[b]);
            if (ev[b]) (void)hipEventDestroy(ev[b]);
        }
    }
    void ensure() {
        for (int b = 0; b < 2; b++) {
            if (!buf[b]) hip_check(hipHostMalloc(&buf[b], CHUNK, hipHostMallocDefault), "pinned staging allocation");
            if (!ev[b]) hip_check(hipEventCreateWithFlags(&ev[b], hipEventDisableTiming), "event");
        }
    }
    // Copies `bytes` from device memory to host memory; the stream's earlier work is
    // ordered before the copy.  Returns after the data is in `dst`.
    void download(void* dst, const void* src, size_t bytes, hipStream_t stream) {
        ensure();
        const size_t n = (bytes + CHUNK - 1) / CHUNK;
        auto drain = [&](size_t k) {
            const int b = int(k & 1);
            hip_check(hipEventSynchronize(ev[b]), "staging sync");
            const size_t off = k * CHUNK, len = std::min(CHUNK, bytes - off);
            std::memcpy(static_cast<char*>(dst) + off, buf[b], len);
        };
        for (size_t k = 0; k < n; k++) {
            const int b = int(k & 1);
            if (k >= 2) drain(k - 2);
            const size_t off = k * CHUNK, len = std::min(CHUNK, bytes - off);
            hip_check(hipMemcpyAsync(buf[b], static_cast<const char*>(src) + off, len, hipMemcpyDeviceToHost, stream),
                      "staged copy");
            hip_check(hipEventRecord(ev[b], stream), "event record");
        }
        for (size_t k = n >= 2 ? n - 2 : 0; k < n; k++) drain(k);
    }
};

const Runner* runner_for(int prec, int lprec) {
    switch (prec) {
    case LBM_F64: return runner_f64();
    case LBM_F32: return runner_f32();
    case LBM_F16_SHIFTED: return lprec == LBM_F16_SHIFTED ? runner_f16s_l16() : runner_f16s();
    default: return runner_bf16s();
    }
}

// Sets the cotangent scale record: the field is (about to be) written with exponent
// e_cur, and the next adjoint step writes with e_next.
__global__ void k_lam_init(int* lsc, int e_cur, int e_next) {
    if (threadIdx.x == 0) {
        lsc[0] = e_cur;
        lsc[1] = e_next;
        lsc[2] = 0;
        lsc[3] = 0;
    }
    if (threadIdx.x < LAM_SLOTS) lsc[LAM_SLOT0 + threadIdx.x] = 0;
}

size_t elem_bytes(int prec) { return prec == LBM_F64 ? 8 : (prec == LBM_F32 ? 4 : 2); }

// Lattice tables on the host (same order as lattice.h).
int lat_c(int Q, int q, int d) { return Q == 19 ? cvec<19>(q, d) : cvec<27>(q, d); }
double lat_w(int Q, int q) { return Q == 19 ? weight<19>(q) : weight<27>(q); }

// Reverse steps whose alpha_bar terms one FP32 partial gathers before the fold into
// FP64: the partial's rounding stays near 32 ulp of FP32 per run, far below the
// reduced-precision paths' own error, and the fold costs one FP64 read and write a run.
constexpr uint64_t ABAR_RUN = 32;

// Exact rationals for the known-stress operator of the regularized ports: num/den in
// lowest terms, den > 0; every operation refuses a result beyond 62 bits.
struct Ratio {
    long long num = 0, den = 1;
    static Ratio of(__int128 n, __int128 d) {
        if (d < 0) n = -n, d = -d;
        __int128 a = n < 0 ? -n : n, b = d;
        while (b != 0) {
            const __int128 t = a % b;
            a = b;
            b = t;
        }
        if (a > 1) n /= a, d /= a;
        const __int128 lim = __int128(1) << 62;
        if (d == 0 || n >= lim || -n >= lim || d >= lim) fail(LBM_ERR_INVALID, "stress operator out of range");
        return Ratio{(long long)n, (long long)d};
    }
    Ratio operator+(const Ratio& o) const { return of(__int128(num) * o.den + __int128(o.num) * den, __int128(den) * o.den); }
    Ratio operator*(const Ratio& o) const { return of(__int128(num) * o.num, __int128(den) * o.den); }
};

// known_stress_operator (ports.rs:409-440) of D3Q19 face `face` as op[r * 14 + k]: the
// pseudo-inverse of the known rows (c_axis != sign, ascending q) of the stress basis
// 4.5 w_q (Q00, Q11, Q22, 2 Q01, 2 Q02, 2 Q12) weighted by 1/sqrt(w), times 1/sqrt(w)
// per column.  The known rows B have rank 6, so that is (B^T W^-1 B)^-1 B^T W^-1, a
// matrix of rationals: it is formed exactly and each entry rounded once to the nearest
// double (the reference's SVD route agrees to a few units in the last place).
void known_stress_operator(int face, double* op) {
    const int axis = face / 2, sign = face % 2 == 0 ? 1 : -1;
    static const int ia[6] = {0, 1, 2, 0, 0, 1}, ib[6] = {0, 1, 2, 1, 2, 2};
    std::vector<int> known;
    for (int q = 0; q < 19; q++)
        if (cvec<19>(q, axis) != sign) known.push_back(q);
    if (known.size() != 14) fail(LBM_ERR_INVALID, "stress fit needs 14 known directions");
    Ratio b[14][6], winv[14];
    for (int k = 0; k < 14; k++) {
        const int q = known[k], order = vorder<19>(q);
        winv[k] = Ratio{order == 0 ? 3 : (order == 1 ? 18 : 36), 1};
        const Ratio scale = order == 0 ? Ratio{3, 2} : (order == 1 ? Ratio{1, 4} : Ratio{1, 8});  // 4.5 w_q
        for (int r = 0; r < 6; r++) {
            const int ca = cvec<19>(q, ia[r]), cb = cvec<19>(q, ib[r]);
            const Ratio e = r < 3 ? Ratio::of(3 * ca * ca - 1, 3) : Ratio{2 * ca * cb, 1};
            b[k][r] = scale * e;
        }
    }
    Ratio m[6][12];  // [N | I] -> [I | N^-1], N = B^T W^-1 B
    for (int r = 0; r < 6; r++)
        for (int c = 0; c < 12; c++) {
            Ratio acc{c == r + 6 ? 1 : 0, 1};
            if (c < 6)
                for (int k = 0; k < 14; k++) acc = acc + b[k][r] * b[k][c] * winv[k];
            m[r][c] = acc;
        }
    for (int c = 0; c < 6; c++) {
        int piv = c;
        while (piv < 6 && m[piv][c].num == 0) piv++;
        if (piv == 6) fail(LBM_ERR_INVALID, "Regularized port stress fit is rank deficient or ill conditioned");
        std::swap(m[c], m[piv]);
        const Ratio inv = Ratio::of(m[c][c].den, m[c][c].num);
        for (int j = 0; j < 12; j++) m[c][j] = m[c][j] * inv;
        for (int r = 0; r < 6; r++) {
            if (r == c || m[r][c].num == 0) continue;
            const Ratio f{-m[r][c].num, m[r][c].den};
            for (int j = 0; j < 12; j++) m[r][j] = m[r][j] + f * m[c][j];
        }
    }
    for (int r = 0; r < 6; r++)
        for (int k = 0; k < 14; k++) {
            Ratio acc{0, 1};
            for (int c = 0; c < 6; c++) acc = acc + m[r][6 + c] * b[k][c];
            acc = acc * winv[k];
            if (acc.num >= (1LL << 53) || -acc.num >= (1LL << 53) || acc.den >= (1LL << 53))
                fail(LBM_ERR_INVALID, "stress operator entry beyond exact doubles");
            op[r * 14 + k] = double(acc.num) / double(acc.den);  // one rounding: both exact
        }
}

// Weighted Gram-Schmidt moment basis (moving/collision.rs:87-153): monomials up to
// degree 6 in the order mass, momentum, bulk, shear, then degrees 3..6; group codes
// 0 mass, 1 momentum, 2 bulk, 3 shear, 4 odd, 5 even-high.
void mrt_basis(int Q, std::vector<double>& rows, std::vector<double>& norms, std::vector<int>& groups) {
    auto ipow = [](double v, int e) {
        double r = 1.0;
        for (int k = 0; k < e; k++) r *= v;
        return r;
    };
    std::vector<std::pair<std::vector<double>, int>> cand;
    auto poly = [&](auto fn, int group) {
        std::vector<double> p(Q);
        for (int i = 0; i < Q; i++) {
            const double v[3] = {double(lat_c(Q, i, 0)), double(lat_c(Q, i, 1)), double(lat_c(Q, i, 2))};
            p[i] = fn(v);
        }
        cand.emplace_back(std::move(p), group);
    };
    poly([](const double*) { return 1.0; }, 0);
    poly([](const double* v) { return v[0]; }, 1);
    poly([](const double* v) { return v[1]; }, 1);
    poly([](const double* v) { return v[2]; }, 1);
    poly([](const double* v) { return v[0] * v[0] + v[1] * v[1] + v[2] * v[2]; }, 2);
    poly([](const double* v) { return 2.0 * v[0] * v[0] - v[1] * v[1] - v[2] * v[2]; }, 3);
    poly([](const double* v) { return v[1] * v[1] - v[2] * v[2]; }, 3);
    poly([](const double* v) { return v[0] * v[1]; }, 3);
    poly([](const double* v) { return v[0] * v[2]; }, 3);
    poly([](const double* v) { return v[1] * v[2]; }, 3);
    for (int degree = 3; degree <= 6; degree++)
        for (int a = 0; a <= 2; a++)
            for (int b = 0; b <= 2; b++) {
                const int c = degree - a - b;
                if (c < 0 || c > 2) continue;
                poly([&](const double* v) { return ipow(v[0], a) * ipow(v[1], b) * ipow(v[2], c); },
                     degree % 2 == 1 ? 4 : 5);
            }
    auto inner = [&](const std::vector<double>& x, const std::vector<double>& y) {
        double s = 0.0;
        for (int i = 0; i < Q; i++) s += lat_w(Q, i) * x[i] * y[i];
        return s;
    };
    rows.clear();
    norms.clear();
    groups.clear();
    std::vector<std::vector<double>> basis;
    for (auto& [p, group] : cand) {
        if (int(basis.size()) == Q) break;
        const double scale = inner(p, p);
        if (scale <= 0.0) continue;
        std::vector<double> v = p;
        for (size_t k = 0; k < basis.size(); k++) {
            const double coef = inner(p, basis[k]) / norms[k];
            for (int i = 0; i < Q; i++) v[i] -= coef * basis[k][i];
        }
        const double n = inner(v, v);
        if (n <= 1e-10 * scale) continue;
        norms.push_back(n);
        groups.push_back(group);
        basis.push_back(v);
    }
    if (int(basis.size()) != Q) fail(LBM_ERR_CONTRACT, "MRT moment basis does not span the lattice");
    for (auto& r : basis) rows.insert(rows.end(), r.begin(), r.end());
}

double odd_rate(double omega, double magic) {
    const double tp = 1.0 / omega - 0.5;
    return 1.0 / (0.5 + magic / tp);
}

}  // namespace

struct lbm_handle {
    int Q = 19, kind = K_TRT, prec = LBM_F32, lprec = LBM_F32;
    bool regularized = false;
    double les_c = 0.0, les_floor = 0.0;  // Smagorinsky closure (K_TRT_LES / K_BGK_LES)
    double wale_c2 = 0.0, wale_floor = 0.0;  // WALE closure (K_TRT_WALE / K_BGK_WALE)
    int device = 0;      // HIP device current at lbm_create (the handle's memory and stream)
    uint64_t every = 0;  // Checkpointing::Every(k); 0 = binomial
    double mach_limit = 0.0, drift_limit = 1e-6;  // admission (0: none)
    int nx = 0, ny = 0, nz = 0;
    int per[3] = {0, 0, 0};
    double omega = 1.0, om = 1.0, tau = 1.0, magic = 3.0 / 16.0, bulk = NAN, even = NAN;
    double a[3] = {0, 0, 0};
    bool has_mask = false;
    Geom g{};
    long long ncell = 0;
    long long arr_elems = 0;  // elements of one population array's share of a field (row-blocked)
    size_t state_bytes = 0, lam_bytes = 0;
    const Runner* run = nullptr;
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    DevBuf F, A, L, alpha, alpha_bar, mask, mrt, pcell, pnb, pkind, pval, pbuf, scal, aos, pool, part, lsc, fold;
    DevBuf stress;            // regularized ports: 6 faces x (6 x 14) known-stress operators
    DevBuf abar32;            // FP32 partial of alpha_bar over a run of reverse steps
    long long pre_off = 0;    // regularized ports: byte offset of the pre-port populations
                              // in every state buffer (after the flow populations), else 0
    DevBuf W;          // working state of a gradient sweep (F keeps state 0 throughout)
    DevBuf ledger;     // mass ledger of the last sweep (grows to the longest horizon)
    bool abar_ready = false;  // alpha_bar holds the result of a gradient or step VJP
    HostStage stage;   // pinned staging for device-to-host copies
    long long npart = 0;
    double alpha_max = 0.0;  // max |alpha| over the design (bounds dissipation seeds)
    std::vector<uint8_t> solid;
    std::vector<int> h_pcell, h_pnb;
    bool st_ports = false;  // some port rows carry a stationary rule (ports.rs)
    uint64_t t = 0;
    uint64_t budget = 0;
    size_t pool_slots = 0;
    int np = 0;
    std::vector<double> h_alpha;
    std::vector<double> h_consts;  // collision constants (the `mrt` array) in FP64
    // Coupled D3Q7 temperature field: its populations follow the Q flow arrays in every
    // state buffer (flow_bytes in), its cotangents the flow cotangents (flow_lam_bytes in).
    bool thermal = false;
    double tkappa = 0.0, tratio = 1.0, tsource = 0.0, tinlet = 0.0;
    double trange = 1.0;  // bound on |T| (storage scale, cotangent bounds)
    float tenc = 1.0f, tdec = 1.0f;  // FP16 storage scale of the temperature populations
    size_t flow_bytes = 0, flow_lam_bytes = 0;
    size_t tlam_elem = 4;  // bytes of one temperature cotangent
    long long n_tsolid = 0;  // cells with phi > 1/2
    DevBuf phi, phi_bar, ubar;  // ubar: velocity cotangent handed from the thermal to the flow adjoint
    DevBuf wu, wgb;             // WALE: u and rho of the state a step reads; gradient cotangents
    // Sponge layers: the specifications (far-field targets are formed from them for every
    // step) and, on the device, the covered fluid cells with their summed strength and parts.
    std::vector<lbm_sponge> sponges;
    int nsc = 0;
    DevBuf sp_cell, sp_sigma, sp_off, sp_part, sp_weight, sp_phase;
    // Occupancy coupling (K_TRT_PSM / K_BGK_PSM): the design array is the occupancy d;
    // solid momentum m (SoA, arithmetic precision), its FP64 cotangent, and a zero design
    // for the Brinkman-velocity outputs (fields).
    int psm_law = -1;  // 0 Psm, 1 PsmSuperposition, 2 Brinkman with a moving solid
    double psm_kappa = 0.0, psm_c_alpha = 0.0, psm_drag_shape = 0.0;
    DevBuf sm, sm_bar, alpha0;
    // Second occupancy endpoint (d1, m1) blended over steps te0..te1, and its cotangents.
    DevBuf d1, sm1, d1_bar, sm1_bar;
    uint64_t te0 = 0, te1 = 1;
    // Subcycle (lbm_subcycle): the k-th of m substeps blends at its midpoint, theta =
    // (k - 1/2) / m (Field::params, moving/field.rs:1049); gacc accumulates the endpoints'
    // impulse shares in the forward sweep, gbar holds their cotangents in the reverse one.
    bool sub_mid = false, gacc_on = false, gbar_on = false;
    DevBuf gacc, gbar;
    // Weight of the second endpoint in the step that reads state t.
    double theta(uint64_t t) const {
        if (!d1.p) return 0.0;
        if (sub_mid) {
            if (t < te0) return 0.0;
            if (t >= te1) return 1.0;
            return (double(t - te0 + 1) - 0.5) / double(te1 - te0);
        }
        if (t <= te0) return 0.0;
        if (t >= te1) return 1.0;
        return double(t - te0) / double(te1 - te0);
    }
    const ThermalRunner* trun = nullptr;

    size_t csize() const { return prec == LBM_F64 ? 8 : 4; }
    // Step grid: one block of b cells (a multiple of 64) along a z row, one row per grid
    // y, one x plane per grid z.  (Blocks of two y rows measured no faster at 256^3.)
    // k_forward and k_adjoint take the row from (blockIdx.z, blockIdx.y) and rely on
    // blockDim.y = 1.
    LaunchCfg cfg() const {
        const int b = std::min(256, ((nz + 63) / 64) * 64);
        return LaunchCfg{dim3((nz + b - 1) / b, ny, nx), dim3(b), stream};
    }
    StepArgs step_args(void* pop, uint64_t step) const {
        StepArgs s{};
        s.g = g;
        s.pop = pop;
        s.pop_out = pop;
        s.alpha = alpha.p;
        s.mask = mask.as<uint32_t>();
        s.omega = omega;
        s.om = om;
        s.tau = tau;
        for (int d = 0; d < 3; d++) s.a[d] = a[d];
        s.mrt = mrt.p;
        s.odd = int(step & 1);
        s.step = uint32_t(step);
        s.wu = wu.p;
        s.sm = sm.p;
        s.d1 = d1.p;
        s.sm1 = sm1.p;
        s.theta = theta(step);
        s.gacc = gacc_on ? gacc.as<double>() : nullptr;
        if (thermal) {  // the flow kernel advances the temperature field in the same pass
            s.tpop = tpop(static_cast<const void*>(pop));
            s.tpop_out = tpop(pop);
            s.phi = phi.p;
            s.kappa = tkappa;
            s.ratio_m1 = tratio - 1.0;
            s.source = tsource;
            s.tenc = tenc;
            s.tdec = tdec;
        }
        s.diag = 0;
        s.obj.kind = OBJ_NONE;
        s.part = nullptr;
        s.npart = npart;
        s.stage = fold.as<double>();
        s.ticket = reinterpret_cast<unsigned*>(fold.as<double>() + FOLD_VALUES * FOLD_MAX_BLOCKS);
        s.jacc = scal.as<double>();
        s.macc = scal.as<double>() + 1;
        s.dstat = reinterpret_cast<unsigned*>(scal.as<double>() + 4);
        return s;
    }
    // Sponge arguments of the step that reads state `step`: the slots it writes (parity
    // step & 1) and the far-field targets at time t = step (field.rs:1015-1036 with unit
    // spacing and step; the signal's ramp as boundary.rs:110-118).
    SpongeArgs sponge_args(uint64_t step) const {
        SpongeArgs s{};
        s.g = g;
        s.n = nsc;
        s.ns = int(sponges.size());
        s.odd = int(step & 1);
        s.cell = sp_cell.as<int>();
        s.sigma = sp_sigma.as<double>();
        s.off = sp_off.as<int>();
        s.part = sp_part.as<int>();
        s.weight = sp_weight.as<double>();
        s.phase = sp_phase.as<double>();
        const double t = double(step);
        for (size_t l = 0; l < sponges.size(); l++) {
            const lbm_sponge& sp = sponges[l];
            double* r = s.target[l];
            r[0] = sp.density;
            for (int d = 0; d < 3; d++) r[1 + d] = sp.velocity[d];
            if (sp.wave) {
                double ramp = 1.0;
                if (sp.wave_ramp > 0.0 && t < sp.wave_ramp) {
                    double x = t / sp.wave_ramp;
                    if (x < 0.0) x = 0.0;
                    ramp = x * x * x * (10.0 - x * 15.0 + x * x * 6.0);
                }
                const double arg = t * (2.0 * M_PI * sp.wave_frequency) + sp.wave_phase;
                const double sw = ramp * std::sin(arg), cw = ramp * std::cos(arg);
                for (int d = 0; d < 3; d++) {
                    r[4 + d] = sw * sp.wave_amplitude[d];
                    r[7 + d] = cw * sp.wave_amplitude[d];
                }
            }
        }
        return s;
    }
    void* tpop(void* buf) const { return static_cast<char*>(buf) + flow_bytes; }
    const void* tpop(const void* buf) const { return static_cast<const char*>(buf) + flow_bytes; }
    void* tlam() const { return static_cast<char*>(L.p) + flow_lam_bytes; }
    // Thermal step arguments: flow and temperature read from src (parity of `step`),
    // temperature written to dst; no stage cost.
    ThermalArgs thermal_args(const void* src, void* dst, uint64_t step) const {
        ThermalArgs t{};
        t.g = g;
        t.pop = src;
        t.tpop = tpop(src);
        t.tpop_out = dst ? tpop(dst) : nullptr;
        t.tlam = L.p ? tlam() : nullptr;
        t.ubar = ubar.p;
        t.alpha = alpha.p;
        t.phi = phi.p;
        t.alpha_bar = alpha_bar.as<double>();
        t.phi_bar = phi_bar.as<double>();
        for (int d = 0; d < 3; d++) t.a[d] = a[d];
        t.kappa = tkappa;
        t.ratio_m1 = tratio - 1.0;
        t.source = tsource;
        t.tenc = tenc;
        t.tdec = tdec;
        t.odd = int(step & 1);
        t.step = uint32_t(step);
        t.obj_w = 0.0;
        t.part = nullptr;
        t.stage = fold.as<double>();
        t.ticket = reinterpret_cast<unsigned*>(fold.as<double>() + FOLD_VALUES * FOLD_MAX_BLOCKS);
        t.jacc = scal.as<double>();
        return t;
    }
    bool lam16() const { return lprec == LBM_F16_SHIFTED; }
    int* lsc_ptr() const { return lam16() ? lsc.as<int>() : nullptr; }
    // Adjoint-step arguments on state buffer pop (parity odd), no stage cost.
    AdjArgs adj_args(const void* pop, int odd) const {
        AdjArgs a{};
        a.g = g;
        a.pop = pop;
        a.lam = L.p;
        a.alpha = alpha.p;
        a.alpha_bar = alpha_bar.as<double>();
        a.mask = mask.as<uint32_t>();
        a.omega = omega;
        a.om = om;
        a.tau = tau;
        for (int d = 0; d < 3; d++) a.a[d] = this->a[d];
        a.mrt = mrt.p;
        a.odd = odd;
        a.sm = sm.p;
        a.sm_bar = sm_bar.as<double>();
        a.d1 = d1.p;
        a.sm1 = sm1.p;
        a.d1_bar = d1_bar.as<double>();
        a.sm1_bar = sm1_bar.as<double>();
        a.gbar = gbar_on ? gbar.as<double>() : nullptr;
        a.obj.kind = OBJ_NONE;
        a.part = part.as<double>();
        a.npart = npart;
        a.stage = fold.as<double>();
        a.ticket = reinterpret_cast<unsigned*>(fold.as<double>() + FOLD_VALUES * FOLD_MAX_BLOCKS);
        a.lsc = lsc_ptr();
        a.s_next = 0.0;
        return a;
    }
    PortArgs port_args(uint64_t step) const {
        PortArgs p{};
        p.g = g;
        p.np = np;
        p.cell = pcell.as<int>();
        p.nb = pnb.as<int>();
        p.kind = pkind.as<int>();
        p.value = pval.as<double>();
        p.mask = mask.as<uint32_t>();
        p.odd = int(step & 1);
        p.st = st_ports ? 1 : 0;
        p.stress = stress.as<double>();
        p.pre_off = pre_off;
        return p;
    }
};

namespace {

// Objective of one state index s (1-based) inside a sweep of n steps, from a template
// holding the plane geometry (built once per call by objective_template).
ObjArgs stage_obj(const ObjArgs& tmpl, const lbm_objective* o, uint64_t s, uint64_t n, bool with_terminal) {
    ObjArgs r = tmpl;
    r.kind = OBJ_NONE;
    if (!o || o->kind == LBM_OBJ_NONE) return r;
    double w = 0.0;
    if (s >= o->t_begin && s <= o->t_end && s >= 1) w += o->weight;
    if (with_terminal && s == n) w += o->terminal_weight;
    if (w == 0.0) return r;
    if (o->kind == LBM_OBJ_THERMAL) {
        // temperature part on the thermal kernels, pressure-drop part on the flow kernels
        r.tw = w * o->vector_weights[0] * tmpl.inv_ns;
        r.weight = w * o->vector_weights[1];
        r.kind = r.weight != 0.0 ? int(OBJ_PRESSURE_DROP) : int(OBJ_NONE);
        return r;
    }
    r.kind = int(o->kind);
    r.weight = w;
    return r;
}

long long fluid_cells(const lbm_handle* h) {
    long long n = h->ncell;
    if (!h->solid.empty())
        for (long long x = 0; x < h->ncell; x++) n -= h->solid[x] ? 1 : 0;
    return n;
}

// Per-cell weights of w . vector (vector_vjp's weights, solver.rs:905-912) without the
// design-volume term.
void vector_weights(const lbm_handle* h, const double w[6], const lbm_vector_scales& sc, ObjArgs& r) {
    const double fc = double(fluid_cells(h));
    r.vw[0] = w[0] / fc;
    r.vw[1] = w[1] / fc;
    r.vw[2] = w[2] / fc;
    r.vw[3] = 0.5 * w[3] * sc.cell_volume;
    r.vw[4] = w[4] * sc.cell_volume;
    r.vel = sc.velocity;
    r.dens = sc.density;
    r.inv_dt = 1.0 / sc.step_s;
}

ObjArgs objective_template(const lbm_handle* h, const lbm_objective* o) {
    ObjArgs r{};
    r.kind = OBJ_NONE;
    if (!o || o->kind == LBM_OBJ_NONE) return r;
    if (is_psm(h->kind) && o->kind != LBM_OBJ_PRESSURE_DROP && o->kind != LBM_OBJ_SOLID_FORCE)
        fail(LBM_ERR_CONTRACT, "with the occupancy coupling the objective is the pressure drop or the solid force "
                               "(the other objectives read a Brinkman coefficient)");
    if (o->kind == LBM_OBJ_SOLID_FORCE) {
        if (!is_psm(h->kind)) fail(LBM_ERR_CONTRACT, "the solid-force objective needs the occupancy coupling");
        if (o->terminal_weight != 0.0)
            fail(LBM_ERR_CONTRACT, "the solid-force objective is a stage cost (terminal_weight must be 0)");
        for (int k = 0; k < 3; k++) {
            if (!std::isfinite(o->vector_weights[k])) fail(LBM_ERR_CONTRACT, "solid-force weights must be finite");
            r.vw[k] = o->vector_weights[k];
        }
    }
    if (o->kind == LBM_OBJ_VECTOR) vector_weights(h, o->vector_weights, o->scales, r);
    r.axis = int(o->axis);
    const int len = o->axis == 0 ? h->nx : (o->axis == 1 ? h->ny : h->nz);
    r.plane_a = ((o->plane_a % len) + len) % len;
    r.plane_b = ((o->plane_b % len) + len) % len;
    if (o->kind == LBM_OBJ_THERMAL && o->vector_weights[0] != 0.0) {
        if (h->n_tsolid == 0) fail(LBM_ERR_CONTRACT, "no cell has thermal design phi > 1/2 (mean solid temperature)");
        r.inv_ns = 1.0 / double(h->n_tsolid);
    }
    if (o->kind == LBM_OBJ_PRESSURE_DROP || o->kind == LBM_OBJ_THERMAL) {
        long long na = 0, nb = 0;
        const long long plane = h->ncell / len;
        if (h->solid.empty()) {
            na = nb = plane;
        } else {
            for (long long x = 0; x < h->ncell; x++) {
                if (h->solid[x]) continue;
                const int k = int(x % h->nz), j = int((x / h->nz) % h->ny), i = int(x / ((long long)h->nz * h->ny));
                const int c = o->axis == 0 ? i : (o->axis == 1 ? j : k);
                na += c == r.plane_a;
                nb += c == r.plane_b;
            }
        }
        if (na == 0 || nb == 0) fail(LBM_ERR_CONTRACT, "pressure-drop plane has no fluid cells");
        r.inv_na = 1.0 / double(na);
        r.inv_nb = 1.0 / double(nb);
    }
    return r;
}

void validate_objective(const lbm_handle* h, const lbm_objective* o, uint64_t n) {
    if (!o || o->kind == LBM_OBJ_NONE) return;
    if (o->kind > LBM_OBJ_SOLID_FORCE) fail(LBM_ERR_CONTRACT, "unknown objective kind");
    // the exchange of a state is formed by the step that collides it: the final state has none
    if (o->kind == LBM_OBJ_SOLID_FORCE && o->weight != 0.0 && o->t_end >= n)
        fail(LBM_ERR_CONTRACT, "the solid-force window must end before the final state (t_end < nsteps)");
    if (o->kind == LBM_OBJ_THERMAL) {
        if (!h->thermal) fail(LBM_ERR_CONTRACT, "the thermal objective needs the coupled temperature field");
        if (!std::isfinite(o->vector_weights[0]) || !std::isfinite(o->vector_weights[1]))
            fail(LBM_ERR_CONTRACT, "thermal objective weights must be finite");
    }
    if (o->kind == LBM_OBJ_VECTOR) {
        for (int k = 0; k < 6; k++)
            if (!std::isfinite(o->vector_weights[k])) fail(LBM_ERR_CONTRACT, "vector weights must be finite");
        const lbm_vector_scales& sc = o->scales;
        if (!(sc.velocity > 0.0) || !(sc.density > 0.0) || !(sc.step_s > 0.0) || !(sc.cell_volume > 0.0) ||
            !std::isfinite(sc.velocity + sc.density + sc.step_s + sc.cell_volume))
            fail(LBM_ERR_CONTRACT, "vector scales must be positive and finite");
    }
    if (o->axis > 2) fail(LBM_ERR_CONTRACT, "objective axis must be 0, 1 or 2");
    if (!std::isfinite(o->weight) || !std::isfinite(o->terminal_weight))
        fail(LBM_ERR_CONTRACT, "objective weights must be finite");
    if (o->weight != 0.0 && (o->t_begin < 1 || o->t_end > n || o->t_begin > o->t_end))
        fail(LBM_ERR_CONTRACT, "objective window must satisfy 1 <= t_begin <= t_end <= nsteps");
    (void)h;
}

void ensure_aos(lbm_handle* h) { h->aos.alloc(size_t(h->ncell) * h->Q * sizeof(double), "layout conversion"); }

// Largest |d phi / d f_q| one stage cost can add to a cell's cotangent (0 for none).
// Pressure drop: weight / (3 * plane cells).  Porous dissipation rho alpha |u|^2:
// |rho_bar| + sum_d |j_bar_d| <= weight * alpha * (3 / rho + 2 sqrt(3) |u|) <= 16 weight
// alpha_max for rho >= 1/2 and |u| <= 1; the FP16 exponent keeps 4x headroom above it.
double seed_bound_flow(const lbm_handle* h, const ObjArgs& o) {
    if (o.kind == OBJ_PRESSURE_DROP) return std::fabs(o.weight) * std::max(o.inv_na, o.inv_nb) / 3.0;
    // |d (w . dp) / d f_q| <= 4 |w|_1 for rho >= 1/2 and |u| <= 1 (dp = sum_i Omega_s,i c_i)
    if (o.kind == OBJ_SOLID_FORCE) return 4.0 * std::fabs(o.weight) * (std::fabs(o.vw[0]) + std::fabs(o.vw[1]) + std::fabs(o.vw[2]));
    if (o.kind == OBJ_DISSIPATION) return 16.0 * std::fabs(o.weight) * h->alpha_max;
    if (o.kind == OBJ_VECTOR) {
        // phi = sum_d A_d u_d + B rho |u|^2: |d phi/d f_q| <= 80 |B| + 8 sum |A_d| for rho >= 1/2, |u| <= 1
        const double v2 = o.vel * o.vel * o.dens;
        const double B = v2 * (std::fabs(o.vw[3]) + std::fabs(o.vw[4]) * h->alpha_max * o.inv_dt);
        const double A = o.vel * (std::fabs(o.vw[0]) + std::fabs(o.vw[1]) + std::fabs(o.vw[2]));
        return std::fabs(o.weight) * (80.0 * B + 8.0 * A);
    }
    return 0.0;
}

// With the temperature field, a stage cost tw * T reaches the flow cotangent in the
// next transposed step through the advecting velocity: |u_bar| <= omega_T |T| sum_i
// 4 w_i |c_i| |lambda_g| <= 2 |T| tw and |d u / d f_q| <= 3 / (rho den), so
// 16 * (bound on |T|) * tw bounds it for rho >= 1/2 (the FP16 exponent keeps 28
// binades below the bound, so the estimate may be generous).
double seed_bound(const lbm_handle* h, const ObjArgs& o) {
    double b = seed_bound_flow(h, o);
    if (h->thermal && o.tw != 0.0) b += 16.0 * h->trange * std::fabs(o.tw);
    return b;
}

// Cotangent scale record before a field of |lambda| <= bound is written (seed or
// upload); s_next bounds the stage seed of the first adjoint step after it.
void lam_init(lbm_handle* h, double bound, double s_next) {
    if (!h->lam16()) return;
    hipLaunchKernelGGL(k_lam_init, dim3(1), dim3(64), 0, h->stream, h->lsc.as<int>(), lam_exponent(bound),
                       lam_exponent(LAM_GROWTH * bound + s_next));
    hip_check(hipGetLastError(), "scale init launch");
}

// Fails when an FP16 cotangent step left the FP16 range (the record counts them).
void check_lam_range(lbm_handle* h) {
    if (!h->lam16()) return;
    int rec[4];
    hip_check(hipMemcpyAsync(rec, h->lsc.p, sizeof(rec), hipMemcpyDeviceToHost, h->stream), "copy");
    hip_check(hipStreamSynchronize(h->stream), "sync");
    if (rec[3] != 0)
        fail(LBM_ERR_CONVERGENCE, std::to_string(rec[3]) +
                                      " adjoint step(s) exceeded the FP16 cotangent range; use FP32 cotangents "
                                      "(adjoint_storage) for this problem");
}

void ensure_adjoint_buffers(lbm_handle* h) {
    if (h->Q == 27 && h->prec == LBM_F64 && (h->kind == K_TRT_PSM || h->kind == K_BGK_PSM))
        fail(LBM_ERR_CONTRACT, "the FP64 D3Q27 Psm adjoint is switched off (a GPU memory fault with no established "
                               "cause); use D3Q19, FP32 or FP16 storage, or the PsmSuperposition or Brinkman law");
    if (h->lam16() && !h->lsc.p) {
        h->lsc.alloc(LAM_REC_INTS * sizeof(int), "cotangent scale");
        hip_check(hipMemsetAsync(h->lsc.p, 0, h->lsc.bytes, h->stream), "memset");
    }
    h->L.alloc(h->lam_bytes, "cotangent state");
    h->A.alloc(h->state_bytes, "ahead state");
    if (h->W.bytes != h->state_bytes) {
        h->W.alloc(h->state_bytes, "working state");
        hip_check(hipMemsetAsync(h->W.p, 0, h->W.bytes, h->stream), "memset");
    }
    h->alpha_bar.alloc(size_t(h->ncell) * sizeof(double), "alpha_bar");
    // FP32 arithmetic: the gradient driver gathers alpha_bar over runs of reverse steps in
    // an FP32 partial (8 bytes a cell a step instead of the FP64 read and write's 16)
    if (h->prec != LBM_F64) h->abar32.alloc(size_t(h->ncell) * sizeof(float), "alpha_bar partial");
    if (h->thermal) {
        h->phi_bar.alloc(size_t(h->ncell) * sizeof(double), "phi_bar");
        h->ubar.alloc(size_t(h->ncell) * 3 * h->csize(), "velocity cotangent");
    }
    if (is_wale(h->kind)) h->wgb.alloc(size_t(h->ncell) * 9 * h->csize(), "WALE gradient cotangents");
    if (is_psm(h->kind)) h->sm_bar.alloc(size_t(h->ncell) * 3 * sizeof(double), "solid momentum cotangent");
    if (h->d1.p) {
        h->d1_bar.alloc(size_t(h->ncell) * sizeof(double), "second occupancy cotangent");
        h->sm1_bar.alloc(size_t(h->ncell) * 3 * sizeof(double), "second solid momentum cotangent");
    }
}

// Snapshot slots for a horizon: min(budget / state bytes, nsteps); the pool only grows.
uint64_t reserve_pool(lbm_handle* h, uint64_t nsteps) {
    ensure_adjoint_buffers(h);
    size_t free_b = 0, total_b = 0;
    hip_check(hipMemGetInfo(&free_b, &total_b), "memory info");
    const uint64_t budget = h->budget ? h->budget : uint64_t(free_b / 2);
    const uint64_t slots = std::min<uint64_t>(budget / h->state_bytes, nsteps);
    if (slots == 0) fail(LBM_ERR_MEMORY, "checkpoint budget holds no snapshot of the state");
    if (h->pool_slots < slots) {
        h->pool.reset();
        h->pool_slots = 0;
        h->pool.alloc(size_t(slots) * h->state_bytes, "revolve snapshot pool");
        h->pool_slots = slots;
    }
    return slots;
}

// Fluid mass of every state of a sweep and the mass the ports and sponges exchange in
// every step (one partial per port or sponge block), for the admission drift
// |M_s - M_0 - sum_{r<s} X_r| / M_0 (solver.rs:707-793).
struct MassLedger {
    DevBuf* buf = nullptr;
    uint64_t n = 0;
    long long nbp = 0;  // port blocks, then sponge blocks
    long long npb = 0;  // port blocks
    size_t used = 0;    // bytes of the ledger this sweep uses
    void init(lbm_handle* h, uint64_t nsteps) {
        n = nsteps;
        npb = (h->np + PORT_BLOCK - 1) / PORT_BLOCK;
        nbp = npb + (h->nsc + SPONGE_BLOCK - 1) / SPONGE_BLOCK;
        used = ((n + 2) + (n + 2) * size_t(nbp)) * sizeof(double);
        buf = &h->ledger;
        buf->grow(used, "mass ledger");
        hip_check(hipMemsetAsync(buf->p, 0, used, h->stream), "memset");
    }
    double* mass(uint64_t s) const { return buf->as<double>() + s; }
    double* exch(uint64_t s) const { return buf->as<double>() + (n + 2) + s * size_t(nbp); }
    // Max relative drift |M_s - M_0 - sum_{r<s} X_r| / M_0 over the states 0..upto.  Call
    // after the stream has been synchronised.
    double drift(uint64_t upto) const {
        std::vector<double> v(used / sizeof(double));
        hip_check(hipMemcpy(v.data(), buf->p, used, hipMemcpyDeviceToHost), "copy");
        const double* M = v.data();
        double ex = 0.0, worst = 0.0;
        for (uint64_t s = 0; s <= std::min(upto, n); s++) {
            worst = std::max(worst, std::fabs(M[s] - M[0] - ex));
            const double* X = v.data() + (n + 2) + s * size_t(nbp);
            for (long long b = 0; b < nbp; b++) ex += X[b];
        }
        return worst / M[0];
    }
    double drift() const { return drift(n); }
};

// Diagnostics of a sweep's states as the fold kernels left them in the scalar record:
// [0] max |u|^2 (float bits), [1] bad cells, [2] order-preserving key of max -f.
struct SweepDiag {
    double max_mach = 0.0, min_population = NAN;
    uint64_t bad_cells = 0;
};

SweepDiag read_diag(lbm_handle* h) {
    unsigned ds[3];
    hip_check(hipMemcpyAsync(ds, h->scal.as<double>() + 4, sizeof(ds), hipMemcpyDeviceToHost, h->stream), "copy");
    hip_check(hipStreamSynchronize(h->stream), "sync");
    SweepDiag d;
    float umax;
    std::memcpy(&umax, &ds[0], 4);
    d.max_mach = std::sqrt(3.0) * std::sqrt(double(umax));
    d.bad_cells = ds[1];
    if (ds[2] != 0) {  // inverse of float_key
        const unsigned u = (ds[2] & 0x80000000u) ? (ds[2] & 0x7fffffffu) : ~ds[2];
        float fneg;
        std::memcpy(&fneg, &u, 4);
        d.min_population = -double(fneg);
    }
    return d;
}

// LbmProblem::admit (solver.rs:707-793): refuse a trajectory with a non-finite or
// non-positive fluid population, a Mach number above the limit or a relative mass drift
// above the bound.  No-op when the handle has no Mach limit.
void admit(const lbm_handle* h, const SweepDiag& d, double drift) {
    if (h->mach_limit <= 0.0) return;
    if (d.bad_cells != 0)
        fail(LBM_ERR_CONVERGENCE, "LBM trajectory has nonpositive or nonfinite populations (" +
                                      std::to_string(d.bad_cells) + " cell states)");
    if (d.max_mach > h->mach_limit || !(drift <= h->drift_limit)) {
        char msg[160];
        std::snprintf(msg, sizeof(msg), "LBM trajectory rejected: Mach=%.17g, mass drift=%.17g", d.max_mach, drift);
        fail(LBM_ERR_CONVERGENCE, msg);
    }
}

// The flow's transposed step on the cotangent buffer (a.pop and a.odd name the state):
// the sponge transpose first; with the WALE closure, the velocity pre-pass before the
// step and the gradient transpose after.
void flow_adjoint(lbm_handle* h, AdjArgs& a, hipEvent_t after = nullptr) {
    if (h->nsc) {  // transpose of the sponge blend on the post-collision cotangents
        SpongeArgs sa = h->sponge_args(0);
        sa.odd = a.odd;
        hip_check(h->run->sponge_vjp(h->Q, sa, a.lam, h->lsc_ptr(), h->stream), "sponge transpose launch");
    }
    if (is_wale(h->kind)) {
        hip_check(h->run->wale_u(h->Q, h->has_mask, h->g, a.pop, h->mask.as<uint32_t>(), a.odd, h->wu.p, h->cfg()),
                  "WALE velocity launch");
        a.wu = h->wu.p;
        a.wgb = h->wgb.p;
    }
    hip_check(h->run->adjoint(h->Q, h->kind, h->has_mask, a, h->cfg(), after), "adjoint launch");
    if (is_wale(h->kind))
        hip_check(h->run->wale_transpose(h->Q, h->has_mask, h->g, h->mask.as<uint32_t>(), h->wu.p, h->wgb.p, a.lam,
                                         a.odd, h->cfg()),
                  "WALE transpose launch");
}

// One forward step of state `step` read from `src`, written to `dst`; with a ledger,
// the step records the mass of the state it reads and the port exchange it makes.
void step_into(lbm_handle* h, void* src, void* dst, uint64_t step, const ObjArgs& obj, const MassLedger* led,
               uint64_t s_index, const std::pair<hipEvent_t, hipEvent_t>* kspan = nullptr) {
    const bool diag = led != nullptr;
    StepArgs s = h->step_args(src, step);
    s.pop_out = dst;
    s.obj = obj;
    s.diag = diag ? 1 : 0;
    if (h->thermal) {
        s.tpop_out = h->tpop(dst);
        s.tw = obj.tw;
    }
    if (obj.kind != OBJ_NONE || obj.tw != 0.0 || diag) s.part = h->part.as<double>();
    if (diag) s.macc = led->mass(s_index);
    if (is_wale(h->kind))  // velocities of the state read, before the in-place step overwrites them
        hip_check(h->run->wale_u(h->Q, h->has_mask, h->g, src, h->mask.as<uint32_t>(), int(step & 1), h->wu.p,
                                 h->cfg()),
                  "WALE velocity launch");
    if (kspan) hip_check(hipEventRecord(kspan->first, h->stream), "event record");
    hip_check(h->run->forward(h->Q, h->kind, h->has_mask, s, h->cfg(), kspan ? kspan->second : nullptr),
              "forward step launch");
    if (h->nsc)  // blend of the sponge cells' post-collision values, before the ports
        hip_check(h->run->sponge(h->Q, h->sponge_args(step), dst, diag ? led->exch(s_index) + led->npb : nullptr,
                                 h->stream),
                  "sponge launch");
    if (h->np) {
        hip_check(h->run->port_apply(h->Q, h->has_mask, h->port_args(step + 1), dst,
                                     diag ? led->exch(s_index) : nullptr, h->stream),
                  "port launch");
        if (h->thermal)
            hip_check(h->trun->ports(h->port_args(step + 1), h->tpop(dst), h->tinlet, h->tenc, h->tdec,
                                     uint32_t(step), h->stream),
                      "thermal port launch");
    }
}

// Stage cost of the temperature field on a final state (into J) and, with do_seed, its
// cotangent seed lambda_g.
void thermal_final(lbm_handle* h, void* buf, uint64_t step, double obj_w, double seed_w, int do_seed) {
    if (!h->thermal || (obj_w == 0.0 && !do_seed)) return;
    ThermalArgs ta = h->thermal_args(buf, nullptr, step);
    ta.obj_w = obj_w;
    if (obj_w != 0.0) ta.part = h->part.as<double>();
    hip_check(h->trun->seed(ta, seed_w, do_seed, h->cfg()), "thermal objective launch");
}

// Arguments of the objective kernel on the final state of a sweep, with its diagnostics
// (Mach, populations) folded into the sweep's record.  The occupancy couplings measure
// the fluid velocity u = j / rho + a / 2 (zero design), the velocity their collisions
// start from.
StepArgs final_state_args(lbm_handle* h, void* buf, uint64_t step) {
    StepArgs ea = h->step_args(buf, step);
    ea.part = h->part.as<double>();
    ea.diag = 1;
    ea.gacc = nullptr;
    if (is_psm(h->kind)) ea.alpha = h->alpha0.p;
    return ea;
}

void step_once(lbm_handle* h, void* buf, uint64_t step, const ObjArgs& obj, const MassLedger* led,
               uint64_t s_index) {
    step_into(h, buf, buf, step, obj, led, s_index);
}

void download_state(lbm_handle* h, const void* buf, uint64_t step, double* out, int layout) {
    ensure_aos(h);
    hip_check(h->run->get_state(h->Q, h->has_mask, h->g, buf, h->mask.as<uint32_t>(), int(step & 1),
                                h->aos.as<double>(), h->cfg()),
              "get_state launch");
    const size_t n = size_t(h->ncell) * h->Q;
    if (layout == 0) {
        hip_check(hipMemcpyAsync(out, h->aos.p, n * sizeof(double), hipMemcpyDeviceToHost, h->stream), "copy out");
        hip_check(hipStreamSynchronize(h->stream), "sync");
    } else {
        std::vector<double> tmp(n);
        hip_check(hipMemcpyAsync(tmp.data(), h->aos.p, n * sizeof(double), hipMemcpyDeviceToHost, h->stream),
                  "copy out");
        hip_check(hipStreamSynchronize(h->stream), "sync");
        for (long long x = 0; x < h->ncell; x++)
            for (int q = 0; q < h->Q; q++) out[size_t(q) * h->ncell + x] = tmp[size_t(x) * h->Q + q];
    }
}

void upload_state(lbm_handle* h, void* buf, uint64_t step, const double* in, int layout) {
    ensure_aos(h);
    const size_t n = size_t(h->ncell) * h->Q;
    if (layout == 0) {
        hip_check(hipMemcpyAsync(h->aos.p, in, n * sizeof(double), hipMemcpyHostToDevice, h->stream), "copy in");
        hip_check(hipStreamSynchronize(h->stream), "sync");
    } else {
        std::vector<double> tmp(n);
        for (long long x = 0; x < h->ncell; x++)
            for (int q = 0; q < h->Q; q++) tmp[size_t(x) * h->Q + q] = in[size_t(q) * h->ncell + x];
        hip_check(hipMemcpyAsync(h->aos.p, tmp.data(), n * sizeof(double), hipMemcpyHostToDevice, h->stream),
                  "copy in");
        hip_check(hipStreamSynchronize(h->stream), "sync");
    }
    hip_check(h->run->set_state(h->Q, h->has_mask, h->g, buf, h->mask.as<uint32_t>(), int(step & 1),
                                h->aos.as<double>(), h->cfg()),
              "set_state launch");
}

// Sponge layers on the host: the covered fluid cells in ascending order, their summed
// strength and their parts (sponge index, s_l, cos theta, sin theta) in sponge order.
struct SpongeField {
    std::vector<int> cell, off{0}, part;
    std::vector<double> sigma, weight, phase;
};

// ramp of moving/sponge.rs:72-81, in its operation order.
double sponge_ramp(double s) {
    if (s <= 0.0) return 0.0;
    if (s >= 1.0) return 1.0;
    return s * s * s * (10.0 - 15.0 * s + 6.0 * s * s);
}

// SpongeField::new (moving/sponge.rs:112-181) on the lattice with unit spacing and the
// origin at 0 (cell centres at ijk + 1/2), with the reference's contract checks
// (TravellingWave::validate, sponge.rs:33-60); solid cells are not listed.
SpongeField build_sponges(const lbm_handle* h, const lbm_config* cfg) {
    SpongeField f;
    if (cfg->n_sponges == 0) return f;
    static_assert(LBM_MAX_SPONGES == MAX_SPONGES, "sponge capacity");
    if (!cfg->sponges) fail(LBM_ERR_CONTRACT, "n_sponges > 0 needs a sponge array");
    if (cfg->n_sponges > LBM_MAX_SPONGES) fail(LBM_ERR_CONTRACT, "at most 8 sponge layers");
    static const char* names[6] = {"xmin", "xmax", "ymin", "ymax", "zmin", "zmax"};
    const long long n = h->ncell;
    const int N[3] = {h->nx, h->ny, h->nz};
    std::vector<double> sigma(n, 0.0);
    struct Part {
        int l;
        double s, c, sn;
    };
    std::vector<std::vector<Part>> parts;
    std::vector<int> slot(n, -1);
    auto finite3 = [](const double* v) { return std::isfinite(v[0]) && std::isfinite(v[1]) && std::isfinite(v[2]); };
    for (uint32_t l = 0; l < cfg->n_sponges; l++) {
        const lbm_sponge& sp = cfg->sponges[l];
        if (sp.face > LBM_FACE_ZMAX) fail(LBM_ERR_CONTRACT, "unknown sponge face");
        if (sp.wave > 1) fail(LBM_ERR_CONTRACT, "sponge wave must be 0 or 1");
        if (sp.wave) {
            const double* d = sp.wave_direction;
            const double norm = std::sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
            if (!(finite3(sp.wave_amplitude) && finite3(d) && finite3(sp.wave_origin) && std::isfinite(sp.wave_speed) &&
                  sp.wave_speed > 0.0 && std::isfinite(sp.wave_frequency) && sp.wave_frequency > 0.0 &&
                  std::isfinite(sp.wave_phase) && std::isfinite(sp.wave_ramp) && sp.wave_ramp >= 0.0 &&
                  std::fabs(norm - 1.0) <= 1e-9))
                fail(LBM_ERR_CONTRACT, "sponge wave needs finite amplitude and origin, a unit direction, a positive "
                                       "speed and a positive frequency");
        }
        const int axis = int(sp.face) / 2;
        const bool lower = sp.face % 2 == 0;
        if (h->per[axis])
            fail(LBM_ERR_CONTRACT, std::string("sponge on face ") + names[sp.face] + " lies on a periodic axis");
        if (sp.thickness == 0 || 2 * uint64_t(sp.thickness) > uint64_t(N[axis]))
            fail(LBM_ERR_CONTRACT, std::string("sponge on face ") + names[sp.face] +
                                       " must be between 1 cell and half the lattice thick");
        if (!(std::isfinite(sp.strength) && sp.strength > 0.0 && sp.strength < 1.0))
            fail(LBM_ERR_CONTRACT, "sponge strength must lie in (0, 1)");
        if (!std::isfinite(sp.density) || !finite3(sp.velocity))
            fail(LBM_ERR_CONTRACT, "sponge far-field state must be finite");
        const double len = double(sp.thickness);
        for (long long x = 0; x < n; x++) {
            const int co[3] = {int(x / ((long long)h->nz * h->ny)), int((x / h->nz) % h->ny), int(x % h->nz)};
            const int c = co[axis];
            const double depth =
                lower ? double(sp.thickness) - double(c) : double(c) - double(N[axis] - int(sp.thickness)) + 1.0;
            if (depth <= 0.0) continue;
            const double sv = sp.strength * sponge_ramp((depth - 0.5) / len);
            if (sv <= 0.0) continue;
            sigma[x] += sv;
            double pc = 1.0, ps = 0.0;
            if (sp.wave) {
                double dist = 0.0;
                for (int a = 0; a < 3; a++) dist += (double(co[a]) + 0.5 - sp.wave_origin[a]) * sp.wave_direction[a];
                const double theta = 2.0 * M_PI * sp.wave_frequency * dist / sp.wave_speed;
                pc = std::cos(theta);
                ps = std::sin(theta);
            }
            if (slot[x] < 0) {
                slot[x] = int(parts.size());
                parts.emplace_back();
            }
            parts[slot[x]].push_back(Part{int(l), sv, pc, ps});
        }
    }
    for (long long x = 0; x < n; x++)
        if (sigma[x] >= 1.0) fail(LBM_ERR_CONTRACT, "summed sponge strength must stay below one");
    for (long long x = 0; x < n; x++) {
        if (slot[x] < 0 || (!h->solid.empty() && h->solid[x])) continue;
        f.cell.push_back(int(x));
        f.sigma.push_back(sigma[x]);
        for (const Part& p : parts[slot[x]]) {
            f.part.push_back(p.l);
            f.weight.push_back(p.s);
            f.phase.push_back(p.c);
            f.phase.push_back(p.sn);
        }
        f.off.push_back(int(f.part.size()));
    }
    return f;
}

void upload(DevBuf& b, const void* src, size_t bytes, const char* what) {
    b.alloc(bytes, what);
    hip_check(hipMemcpy(b.p, src, bytes, hipMemcpyHostToDevice), what);
}

void reset_scalars(lbm_handle* h) { hip_check(hipMemsetAsync(h->scal.p, 0, h->scal.bytes, h->stream), "memset"); }

float elapsed(lbm_handle* h) {
    float ms = 0.0f;
    hip_check(hipEventElapsedTime(&ms, h->ev0, h->ev1), "event time");
    return ms;
}

}  // namespace

// ---------------------------------------------------------------------------------
// Workspace of lbm_vector_noalloc, for callers that cannot allocate device memory: the
// per-block partial sums of the response kernel and of the design volume live in this
// file-scope device array instead of a per-call allocation.  Every GPU the library runs
// on holds its own copy until the process ends; calls on one GPU share it (and keep the
// last call's contents), so calls that overlap on one GPU race.
// ---------------------------------------------------------------------------------
constexpr size_t NOALLOC_SCRATCH_DOUBLES = size_t(1) << 19;  // 4 MiB
constexpr int NOALLOC_PHI_BLOCKS = 256;
static __device__ __align__(256) double scratch[NOALLOC_SCRATCH_DOUBLES];

namespace {

constexpr int NOALLOC_MAX_DEVICES = 64;
std::mutex g_scratch_mu;
double* g_scratch_ptr[NOALLOC_MAX_DEVICES] = {};

// Device address of `scratch` on device dev (which must be current), looked up once per
// device; host code uses only this address (the array's name is a host-side stand-in).
double* scratch_address(int dev) {
    if (dev < 0 || dev >= NOALLOC_MAX_DEVICES) fail(LBM_ERR_DEVICE, "device index out of range for the workspace");
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            fail(LBM_ERR_DEVICE, std::string("workspace address lookup failed: ") + hipGetErrorString(e));
        g_scratch_ptr[dev] = static_cast<double*>(p);
    }
    return g_scratch_ptr[dev];
}

// Block-wide sum of one value per thread (256 threads), result in thread 0, fixed order.
__device__ double block_sum256(double v) {
    __shared__ double red[4];
    v = wave_sum(v);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
    __syncthreads();
    const double r = red[0] + red[1] + red[2] + red[3];
    __syncthreads();
    return r;
}

// part[b] = sum over x = b*256 + t (+ k * gridDim * 256) of phi[x], fixed order.
__global__ __launch_bounds__(256) void k_noalloc_phi(const double* phi, long long n, double* part) {
    double acc = 0.0;
    const long long step = (long long)gridDim.x * 256;
    for (long long x = (long long)blockIdx.x * 256 + threadIdx.x; x < n; x += step) acc += phi[x];
    acc = block_sum256(acc);
    if (threadIdx.x == 0) part[blockIdx.x] = acc;
}

// One block: folds the response kernel's partials sums[5 b + c] (b < nb) and the volume
// partials (nphi of them) in a fixed order and writes the six responses of
// LbmProblem::vector (solver.rs:874-894).
__global__ __launch_bounds__(256) void k_noalloc_vector(const double* sums, long long nb, const double* phi_part,
                                                        int nphi, double inv_fluid, double cell_volume,
                                                        double* out) {
    double acc[6];
    for (int c = 0; c < 5; c++) {
        double a = 0.0;
        for (long long b = threadIdx.x; b < nb; b += 256) a += sums[5 * b + c];
        acc[c] = block_sum256(a);
    }
    acc[5] = block_sum256(int(threadIdx.x) < nphi ? phi_part[threadIdx.x] : 0.0);
    if (threadIdx.x == 0) {
        out[0] = acc[0] * inv_fluid;
        out[1] = acc[1] * inv_fluid;
        out[2] = acc[2] * inv_fluid;
        out[3] = 0.5 * acc[3] * cell_volume;
        out[4] = acc[4] * cell_volume;
        out[5] = acc[5] * cell_volume;
    }
}

}  // namespace

extern "C" {

int lbm_abi_version(void) { return LBM_ABI_VERSION; }

int lbm_last_error(char* buf, uint64_t len) {
    if (!buf || len == 0) return LBM_ERR_CONTRACT;
    const size_t n = std::min<size_t>(len - 1, g_last_error.size());
    std::memcpy(buf, g_last_error.data(), n);
    buf[n] = 0;
    return LBM_OK;
}

int lbm_create(const lbm_config* cfg, const uint8_t* solid_mask, uint64_t mask_len, lbm_handle** out) {
    return guarded([&] {
        if (!cfg || !out) fail(LBM_ERR_CONTRACT, "null config or output handle");
        *out = nullptr;
        if (cfg->struct_size != sizeof(lbm_config)) fail(LBM_ERR_CONTRACT, "lbm_config struct_size mismatch");
        auto h = std::make_unique<lbm_handle>();
        if (cfg->lattice != 19 && cfg->lattice != 27) fail(LBM_ERR_CONTRACT, "lattice must be D3Q19 or D3Q27");
        h->Q = int(cfg->lattice);
        if (cfg->collision > LBM_REGULARIZED) fail(LBM_ERR_CONTRACT, "unknown collision operator");
        // The regularized operator is the moment-space operator with its own rates.
        h->regularized = cfg->collision == LBM_REGULARIZED;
        h->kind = h->regularized ? int(K_MRT) : int(cfg->collision);
        if (!(cfg->les_constant >= 0.0) || !std::isfinite(cfg->les_constant) || !(cfg->les_norm_floor >= 0.0) ||
            !std::isfinite(cfg->les_norm_floor))
            fail(LBM_ERR_CONTRACT, "Smagorinsky constant and norm floor must be finite and non-negative");
        if (!(cfg->wale_constant >= 0.0) || !std::isfinite(cfg->wale_constant))
            fail(LBM_ERR_CONTRACT, "WALE constant must be finite and non-negative");
        if (cfg->wale_constant > 0.0) {
            if (!(cfg->wale_floor > 0.0) || !std::isfinite(cfg->wale_floor))
                fail(LBM_ERR_CONTRACT, "WALE constant and denominator_floor must be finite and positive");
            if (cfg->les_constant > 0.0) fail(LBM_ERR_CONTRACT, "choose one closure: Smagorinsky or WALE");
            if (h->kind != K_TRT && h->kind != K_BGK)
                fail(LBM_ERR_CONTRACT, "the WALE closure applies to the BGK and TRT operators");
            h->kind = h->kind == K_TRT ? int(K_TRT_WALE) : int(K_BGK_WALE);
            h->wale_c2 = cfg->wale_constant * cfg->wale_constant;
            h->wale_floor = cfg->wale_floor;
        }
        if (cfg->les_constant > 0.0) {
            if (h->kind != K_TRT && h->kind != K_BGK)
                fail(LBM_ERR_CONTRACT, "the Smagorinsky closure applies to the BGK and TRT operators");
            h->kind = h->kind == K_TRT ? int(K_TRT_LES) : int(K_BGK_LES);
            h->les_c = cfg->les_constant;
            h->les_floor = cfg->les_norm_floor;
        }
        if (cfg->coupling > LBM_COUPLING_BRINKMAN_OCCUPANCY) fail(LBM_ERR_CONTRACT, "unknown coupling law");
        if (cfg->coupling != LBM_COUPLING_BRINKMAN) {
            if (h->kind != K_TRT && h->kind != K_BGK)
                fail(LBM_ERR_CONTRACT, "the occupancy coupling applies to the plain BGK and TRT operators");
            if (cfg->thermal) fail(LBM_ERR_CONTRACT, "the occupancy coupling and the temperature field are not combined");
            if (cfg->precision == LBM_BF16_SHIFTED)
                fail(LBM_ERR_CONTRACT, "the occupancy coupling supports FP64, FP32 and FP16 storage");
            if (!(cfg->psm_kappa >= 0.0 && cfg->psm_kappa < 1.0))
                fail(LBM_ERR_CONTRACT, "saturation width kappa must lie in [0, 1)");
            h->psm_law = cfg->coupling == LBM_COUPLING_PSM ? 0 : (cfg->coupling == LBM_COUPLING_PSM_SUPERPOSITION ? 1 : 2);
            static const int kinds[3][2] = {{K_TRT_PSM, K_BGK_PSM}, {K_TRT_PSMS, K_BGK_PSMS}, {K_TRT_BRO, K_BGK_BRO}};
            h->kind = kinds[h->psm_law][h->kind == K_TRT ? 0 : 1];
            if (h->psm_law == 2) {
                if (!(std::isfinite(cfg->psm_c_alpha) && cfg->psm_c_alpha >= 0.0))
                    fail(LBM_ERR_CONTRACT, "drag coefficient c_alpha must be finite and non-negative");
                if (!(std::isfinite(cfg->psm_drag_shape) && cfg->psm_drag_shape > 0.0))
                    fail(LBM_ERR_CONTRACT, "drag shape must be finite and positive");
                h->psm_c_alpha = cfg->psm_c_alpha;
                h->psm_drag_shape = cfg->psm_drag_shape;
            }
            h->psm_kappa = cfg->psm_kappa;
        }
        // Admission (solver.rs:434-437: the Mach limit is positive and at most 0.2).
        if (!(cfg->mach_limit == 0.0 || (cfg->mach_limit > 0.0 && cfg->mach_limit <= 0.2)))
            fail(LBM_ERR_CONTRACT, "LBM mach_limit must be 0 (no admission) or in (0, 0.2]; this is not a "
                                   "compressible gas solver");
        if (!(cfg->drift_limit >= 0.0) || !std::isfinite(cfg->drift_limit))
            fail(LBM_ERR_CONTRACT, "drift_limit must be finite and non-negative");
        h->mach_limit = cfg->mach_limit;
        h->drift_limit = cfg->drift_limit > 0.0 ? cfg->drift_limit : 1e-6;
        if (h->kind == K_CUMULANT && h->Q != 27)
            fail(LBM_ERR_CONTRACT, "cumulant collision requires the D3Q27 velocity set (got D3Q19)");
        if (h->kind == K_BGK_GUO && h->Q != 19) fail(LBM_ERR_CONTRACT, "BGK-Guo is the D3Q19 stationary operator");
        if (cfg->precision > LBM_BF16_SHIFTED) fail(LBM_ERR_CONTRACT, "unknown precision");
        h->prec = int(cfg->precision);
        if (h->prec == LBM_BF16_SHIFTED && (is_wale(h->kind) || h->kind == K_TRT_LES || h->kind == K_BGK_LES))
            fail(LBM_ERR_CONTRACT, "BF16 storage supports the operators without closures (FP64, FP32 or FP16 for "
                                   "the Smagorinsky and WALE closures)");
        switch (cfg->adjoint_storage) {
        case LBM_ADJ_AUTO: h->lprec = h->prec == LBM_F64 ? LBM_F64 : (h->prec == LBM_F16_SHIFTED ? LBM_F16_SHIFTED : LBM_F32); break;
        case LBM_ADJ_F64:
            if (h->prec != LBM_F64) fail(LBM_ERR_CONTRACT, "FP64 cotangents need FP64 storage");
            h->lprec = LBM_F64;
            break;
        case LBM_ADJ_F32:
            if (h->prec == LBM_F64) fail(LBM_ERR_CONTRACT, "FP64 storage keeps FP64 cotangents");
            h->lprec = LBM_F32;
            break;
        case LBM_ADJ_F16_SCALED:
            if (h->prec != LBM_F16_SHIFTED) fail(LBM_ERR_CONTRACT, "FP16 cotangents need FP16-shifted storage");
            h->lprec = LBM_F16_SHIFTED;
            break;
        default: fail(LBM_ERR_CONTRACT, "unknown adjoint storage");
        }
        if (is_wale(h->kind)) {  // the gradient transpose adds into the cotangent: FP32/FP64 only
            if (cfg->adjoint_storage == LBM_ADJ_F16_SCALED)
                fail(LBM_ERR_CONTRACT, "the WALE closure keeps FP32 cotangents");
            if (h->lprec == LBM_F16_SHIFTED) h->lprec = LBM_F32;
            if (cfg->thermal) fail(LBM_ERR_CONTRACT, "the WALE closure and the temperature field are not combined");
        }
        h->thermal = cfg->thermal != 0;
        if (cfg->thermal > 1) fail(LBM_ERR_CONTRACT, "thermal must be 0 or 1");
        if (h->thermal) {
            if (h->kind == K_BGK_GUO) fail(LBM_ERR_CONTRACT, "the temperature field couples to the Brinkman operators, not BGK-Guo");
            if (h->prec == LBM_BF16_SHIFTED) fail(LBM_ERR_CONTRACT, "the temperature field supports FP64, FP32 and FP16 storage");
            // the temperature cotangents stay FP32 (FP64 with FP64 storage) next to the
            // flow's, which may be scaled FP16
            h->tlam_elem = h->prec == LBM_F64 ? 8 : 4;
            if (!(cfg->thermal_kappa > 0.0) || !std::isfinite(cfg->thermal_kappa))
                fail(LBM_ERR_CONTRACT, "thermal_kappa must be finite and positive");
            if (!(cfg->thermal_ratio > 0.0) || !std::isfinite(cfg->thermal_ratio))
                fail(LBM_ERR_CONTRACT, "thermal_ratio must be finite and positive");
            if (!std::isfinite(cfg->thermal_source) || !std::isfinite(cfg->thermal_inlet))
                fail(LBM_ERR_CONTRACT, "thermal source and inlet temperature must be finite");
            if (!(cfg->thermal_range >= 0.0) || !std::isfinite(cfg->thermal_range))
                fail(LBM_ERR_CONTRACT, "thermal_range must be finite and non-negative");
            h->tkappa = cfg->thermal_kappa;
            h->tratio = cfg->thermal_ratio;
            h->tsource = cfg->thermal_source;
            h->tinlet = cfg->thermal_inlet;
            h->trange = cfg->thermal_range > 0.0 ? cfg->thermal_range : 1.0;
            if (h->prec == LBM_F16_SHIFTED) {
                const int e = lam_exponent(cfg->thermal_range > 0.0 ? cfg->thermal_range : 1.0);
                h->tenc = std::ldexp(1.0f, e);
                h->tdec = std::ldexp(1.0f, -e);
            }
            h->trun = h->prec == LBM_F64 ? thermal_runner_f64()
                                         : (h->prec == LBM_F32 ? thermal_runner_f32() : thermal_runner_f16s());
        }
        for (int d = 0; d < 3; d++) {
            if (cfg->shape[d] < 2) fail(LBM_ERR_CONTRACT, "LBM shape must contain three integers >= 2");
            if (cfg->shape[d] > 65535) fail(LBM_ERR_CONTRACT, "lattice extent above 65535");
            h->per[d] = cfg->periodic[d] ? 1 : 0;
            if (!std::isfinite(cfg->accel[d])) fail(LBM_ERR_CONTRACT, "acceleration requires three finite components");
            h->a[d] = cfg->accel[d];
        }
        h->nx = cfg->shape[0];
        h->ny = cfg->shape[1];
        h->nz = cfg->shape[2];
        h->ncell = (long long)h->nx * h->ny * h->nz;
        if (cfg->tau != 0.0) {
            if (!std::isfinite(cfg->tau) || cfg->tau <= 0.0) fail(LBM_ERR_CONTRACT, "tau must be finite and positive");
            h->tau = cfg->tau;
            h->omega = 1.0 / cfg->tau;
        } else {
            if (!std::isfinite(cfg->omega) || cfg->omega <= 0.0)
                fail(LBM_ERR_CONTRACT, "omega must be finite and positive");
            h->omega = cfg->omega;
            h->tau = 1.0 / cfg->omega;
        }
        if (!(h->tau >= 0.51 && h->tau <= 2.0))
            fail(LBM_ERR_CONTRACT, "Choose spacing/time step such that 0.51 <= relaxation time <= 2");
        h->magic = cfg->magic;
        const bool trt_like = h->kind == K_TRT || h->kind == K_MRT || h->kind == K_TRT_LES || h->kind == K_TRT_WALE ||
                              h->kind == K_TRT_PSM || h->kind == K_TRT_PSMS || h->kind == K_TRT_BRO;
        if (trt_like && !(std::isfinite(h->magic) && h->magic > 0.0))
            fail(LBM_ERR_CONTRACT, "TRT magic parameter must be finite and positive");
        h->om = trt_like ? odd_rate(h->omega, h->magic) : h->omega;
        h->bulk = cfg->bulk_rate;
        h->even = cfg->even_rate;
        auto rate_ok = [](double r) { return std::isnan(r) || (std::isfinite(r) && r > 0.0 && r < 2.0); };
        if (!rate_ok(h->bulk) || !rate_ok(h->even)) fail(LBM_ERR_CONTRACT, "relaxation rates must lie in (0, 2)");

        // Solid mask and blocked links.
        if (solid_mask) {
            if (mask_len != uint64_t(h->ncell)) fail(LBM_ERR_SHAPE, "solid mask length must equal the cell count");
            h->solid.assign(solid_mask, solid_mask + h->ncell);
            bool any = false, all = true;
            for (auto s : h->solid) {
                any |= s != 0;
                all &= s != 0;
            }
            if (all) fail(LBM_ERR_CONTRACT, "Need fluid cells");
            h->has_mask = any;
            if (any && h->thermal) fail(LBM_ERR_CONTRACT, "the coupled temperature field has no solid cells");
            if (!any) h->solid.clear();
        }

        // Storage geometry (kernels.h, Addressing): row-blocked fields, one extra storage
        // layer (the esoteric-twist ghost) on each non-periodic x and y axis, rows of SZ = nz
        // rounded up to 64 elements (every array row a whole number of 128-byte lines), and
        // on a non-periodic z axis GY ghost rows per plane holding the rows' +z ghost slots.
        // (With the ghost inside the row, pitch nz + 1, neighbouring rows shared a line that
        // two blocks fetched: 15-22 % more HBM reads than the algorithmic bytes at 256^3.)
        h->g.nx = h->nx;
        h->g.ny = h->ny;
        h->g.nz = h->nz;
        h->g.px = h->per[0];
        h->g.py = h->per[1];
        h->g.pz = h->per[2];
        h->g.SX = h->nx + (h->per[0] ? 0 : 1);
        h->g.SY = h->ny + (h->per[1] ? 0 : 1);
        h->g.SZ = ((h->nz + 63) / 64) * 64;
        h->g.GY = h->per[2] ? 0 : (h->g.SY + h->g.SZ - 1) / h->g.SZ;
        // elements of one array's share of a field: a field of Qn arrays holds Qn times this
        h->arr_elems = (long long)h->g.SX * (h->g.SY + h->g.GY) * h->g.SZ;
        // Kernels address a cell's populations as a row block (64-bit) plus a 32-bit byte
        // offset within its plane.
        if (uint64_t(h->g.SY + h->g.GY) * uint64_t(h->g.SZ) * uint64_t(h->Q) *
                std::max({elem_bytes(h->prec), elem_bytes(h->lprec), h->thermal ? h->tlam_elem : size_t(0)}) >=
            (uint64_t(1) << 31))
            fail(LBM_ERR_CONTRACT, "lattice too large: one x plane of the populations would exceed 2 GiB");
        h->flow_bytes = size_t(h->arr_elems) * h->Q * elem_bytes(h->prec);
        h->flow_lam_bytes = size_t(h->arr_elems) * h->Q * elem_bytes(h->lprec);
        const int TQn = h->thermal ? TQ : 0;
        h->state_bytes = size_t(h->arr_elems) * (h->Q + TQn) * elem_bytes(h->prec);
        h->lam_bytes = h->flow_lam_bytes + size_t(h->arr_elems) * TQn * h->tlam_elem;
        h->run = runner_for(h->prec, h->lprec);

        // Ports (moving/boundary.rs:318-380; ports.rs:107-262 for the stationary rules).
        std::vector<int> pkind;
        std::vector<double> pval;
        std::vector<uint8_t> neighbour_row;  // 1: the row reads its inward neighbour
        std::vector<int> is_port(h->ncell, 0);
        bool has_regularized = false;
        std::vector<double> stress_ops;  // the regularized rules' operators, 6 faces x 84
        for (uint32_t p = 0; p < cfg->n_ports; p++) {
            const lbm_port& port = cfg->ports[p];
            if (port.face > LBM_FACE_ZMAX) fail(LBM_ERR_CONTRACT, "unknown port face");
            if (port.kind > LBM_PORT_PRESSURE_REGULARIZED)
                fail(LBM_ERR_CONTRACT,
                     "Port kind must be velocity or pressure (NEE, neighbour, moment or regularized rule)");
            const bool stationary = port.kind >= LBM_PORT_PRESSURE_NEIGHBOR;
            h->st_ports = h->st_ports || stationary;
            const bool moment = port.kind >= LBM_PORT_VELOCITY_MOMENT;  // the cell is its own source
            const bool regularized = port.kind >= LBM_PORT_VELOCITY_REGULARIZED;
            const bool pressure = port.kind == LBM_PORT_PRESSURE || port.kind == LBM_PORT_PRESSURE_NEIGHBOR ||
                                  port.kind == LBM_PORT_PRESSURE_MOMENT || port.kind == LBM_PORT_PRESSURE_REGULARIZED;
            if (stationary && h->Q != 19) fail(LBM_ERR_CONTRACT, "the stationary port rules (ports.rs) are D3Q19");
            if (regularized && (h->a[0] != 0.0 || h->a[1] != 0.0 || h->a[2] != 0.0))
                fail(LBM_ERR_CONTRACT, "regularized_known_stress ports require zero acceleration; forcing-consistent "
                                       "stress reconstruction is not implemented");
            has_regularized = has_regularized || regularized;
            if (stationary && cfg->thermal)
                fail(LBM_ERR_CONTRACT, "the temperature field takes the NEE velocity and pressure ports");
            const int axis = int(port.face) / 2;
            const int inward = (port.face % 2 == 0) ? 1 : -1;
            const int len = axis == 0 ? h->nx : (axis == 1 ? h->ny : h->nz);
            if (h->per[axis]) fail(LBM_ERR_CONTRACT, "A periodic face cannot be a port");
            if (len < 3) fail(LBM_ERR_CONTRACT, "port needs at least three cells along its axis");
            const int layer = inward > 0 ? 0 : len - 1;
            if (port.mask && port.mask_len != uint64_t(h->ncell))
                fail(LBM_ERR_SHAPE, "port mask needs one entry per cell");
            if (!port.mask && port.mask_len) fail(LBM_ERR_CONTRACT, "port mask_len without a mask");
            std::vector<long long> cells;
            for (long long x = 0; x < h->ncell; x++) {
                const int k = int(x % h->nz), j = int((x / h->nz) % h->ny), i = int(x / ((long long)h->nz * h->ny));
                const int c = axis == 0 ? i : (axis == 1 ? j : k);
                const bool wall = !h->solid.empty() && h->solid[x];
                if (port.mask && port.mask[x]) {
                    if (c != layer) fail(LBM_ERR_CONTRACT, "Port mask must lie on its face");
                    if (wall) fail(LBM_ERR_CONTRACT, "Port cells must be fluid");
                }
                if (c == layer && !wall && (!port.mask || port.mask[x])) cells.push_back(x);
            }
            if (cells.empty()) fail(LBM_ERR_CONTRACT, "port has no cells");
            if (port.profile && pressure) fail(LBM_ERR_CONTRACT, "a profile scales a velocity port's velocity");
            if (port.profile && port.profile_len != cells.size())
                fail(LBM_ERR_SHAPE, "port profile needs one factor per port cell");
            for (int d = 0; d < 3; d++)
                if (!std::isfinite(port.value[d])) fail(LBM_ERR_CONTRACT, "port values must be finite");
            if (pressure && !(port.value[0] > 0.0)) fail(LBM_ERR_CONTRACT, "pressure port density must be positive");
            const int t0 = axis == 0 ? 1 : 0, t1 = axis == 2 ? 1 : 2;
            const int row_kind = !stationary ? int(port.kind)
                                             : ST_PORT_BASE + 6 * int(port.kind - LBM_PORT_PRESSURE_NEIGHBOR) +
                                                   int(port.face);
            for (size_t r = 0; r < cells.size(); r++) {
                const long long x = cells[r];
                if (is_port[x]) fail(LBM_ERR_CONTRACT, "port cells overlap");
                const long long step = axis == 0 ? (long long)h->ny * h->nz : (axis == 1 ? h->nz : 1);
                // the moment rules read the port cell itself (Port::sources, ports.rs:74-82)
                const long long nb = moment ? x : x + inward * step;
                if (!moment && !h->solid.empty() && h->solid[nb])
                    fail(LBM_ERR_CONTRACT, "inward neighbour of a port cell is a wall");
                const double prof = port.profile ? port.profile[r] : 1.0;
                if (!std::isfinite(prof)) fail(LBM_ERR_CONTRACT, "port profile must be finite");
                is_port[x] = 1;
                h->h_pcell.push_back(int(x));
                h->h_pnb.push_back(int(nb));
                neighbour_row.push_back(moment ? 0 : 1);
                pkind.push_back(row_kind);
                switch (port.kind) {
                case LBM_PORT_VELOCITY:
                    for (int d = 0; d < 3; d++) pval.push_back(port.value[d] * prof);
                    break;
                case LBM_PORT_PRESSURE:
                    pval.push_back(port.value[0]);
                    pval.push_back(0.0);
                    pval.push_back(0.0);
                    break;
                case LBM_PORT_PRESSURE_NEIGHBOR:
                    for (int d = 0; d < 3; d++) pval.push_back(port.value[d]);
                    break;
                case LBM_PORT_VELOCITY_MOMENT:  // v0 = u - a/2 (ports.rs:352)
                case LBM_PORT_VELOCITY_REGULARIZED:
                    for (int d = 0; d < 3; d++) pval.push_back(port.value[d] * prof - 0.5 * h->a[d]);
                    break;
                default:
                    pval.push_back(port.value[0]);
                    pval.push_back(port.value[1] - 0.5 * h->a[t0]);
                    pval.push_back(port.value[2] - 0.5 * h->a[t1]);
                    break;
                }
            }
        }
        for (size_t r = 0; r < h->h_pnb.size(); r++)
            if (neighbour_row[r] && is_port[h->h_pnb[r]])
                fail(LBM_ERR_CONTRACT, "the inward neighbour of a port cell is itself a port cell");
        {  // the port transposes add into the neighbours' cotangents, one port cell each
            std::vector<int> nbs;
            for (size_t r = 0; r < h->h_pnb.size(); r++)
                if (neighbour_row[r]) nbs.push_back(h->h_pnb[r]);
            std::sort(nbs.begin(), nbs.end());
            if (std::adjacent_find(nbs.begin(), nbs.end()) != nbs.end())
                fail(LBM_ERR_CONTRACT, "two port cells share an inward neighbour (an axis with ports on both faces "
                                       "needs at least four cells)");
        }
        h->np = int(h->h_pcell.size());
        if (has_regularized) {
            // the transposed regularized rule reads the port cells' pre-port populations,
            // which the port pass overwrites: every state buffer keeps them after its flow
            // populations (the stationary rules take no temperature field)
            std::vector<double> ops(6 * 84);
            for (int f = 0; f < 6; f++) known_stress_operator(f, ops.data() + f * 84);
            stress_ops = ops;
            h->pre_off = (long long)((h->flow_bytes + 255) / 256 * 256);
            h->state_bytes = size_t(h->pre_off) + size_t(h->np) * 19 * elem_bytes(h->prec);
        }

        // Sponge layers (moving/sponge.rs:112-181).
        const SpongeField sf = build_sponges(h.get(), cfg);
        h->sponges.assign(cfg->sponges, cfg->sponges + (sf.cell.empty() ? 0 : cfg->n_sponges));
        h->nsc = int(sf.cell.size());

        // Device allocations.
        hip_check(hipGetDevice(&h->device), "device query");
        hip_check(hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking), "stream create");
        hip_check(hipEventCreate(&h->ev0), "event create");
        hip_check(hipEventCreate(&h->ev1), "event create");
        h->F.alloc(h->state_bytes, "population state");
        hip_check(hipMemsetAsync(h->F.p, 0, h->F.bytes, h->stream), "memset");
        h->alpha.alloc(size_t(h->ncell) * h->csize(), "alpha");
        hip_check(hipMemsetAsync(h->alpha.p, 0, h->alpha.bytes, h->stream), "memset");
        h->h_alpha.assign(h->ncell, 0.0);
        if (is_wale(h->kind)) h->wu.alloc(size_t(h->ncell) * 4 * h->csize(), "WALE velocities");
        if (is_psm(h->kind)) {
            h->sm.alloc(size_t(h->ncell) * 3 * h->csize(), "solid momentum");
            hip_check(hipMemsetAsync(h->sm.p, 0, h->sm.bytes, h->stream), "memset");
            h->alpha0.alloc(size_t(h->ncell) * h->csize(), "zero design");
            hip_check(hipMemsetAsync(h->alpha0.p, 0, h->alpha0.bytes, h->stream), "memset");
        }
        if (h->thermal) {
            h->phi.alloc(size_t(h->ncell) * h->csize(), "thermal design");
            hip_check(hipMemsetAsync(h->phi.p, 0, h->phi.bytes, h->stream), "memset");
        }
        h->scal.alloc(64, "scalars");
        {
            const long long b = std::min(256, ((h->nz + 63) / 64) * 64);
            h->npart = ((h->nz + b - 1) / b) * (long long)h->ny * h->nx;
            h->part.alloc(size_t(PART_STRIDE) * size_t(h->npart) * sizeof(double), "block partials");
            h->fold.alloc((FOLD_VALUES * FOLD_MAX_BLOCKS + 2) * sizeof(double), "fold stage");
            hip_check(hipMemsetAsync(h->fold.p, 0, h->fold.bytes, h->stream), "memset");
        }
        if (h->has_mask) {
            std::vector<uint32_t> m(h->ncell, 0);
            for (long long x = 0; x < h->ncell; x++) {
                if (h->solid[x]) {
                    m[x] = SOLID_BIT;
                    continue;
                }
                const int k = int(x % h->nz), j = int((x / h->nz) % h->ny), i = int(x / ((long long)h->nz * h->ny));
                uint32_t bits = 0;
                for (int q = 1; q < h->Q; q++) {
                    const int c[3] = {lat_c(h->Q, q, 0), lat_c(h->Q, q, 1), lat_c(h->Q, q, 2)};
                    int y[3] = {i - c[0], j - c[1], k - c[2]};
                    const int n[3] = {h->nx, h->ny, h->nz};
                    bool blocked = false;
                    for (int d = 0; d < 3; d++) {
                        if (y[d] < 0 || y[d] >= n[d]) {
                            if (!h->per[d]) blocked = true;
                            y[d] = ((y[d] % n[d]) + n[d]) % n[d];
                        }
                    }
                    const long long yx = ((long long)y[0] * h->ny + y[1]) * h->nz + y[2];
                    if (blocked || h->solid[yx]) bits |= 1u << q;
                }
                m[x] = bits;
            }
            h->mask.alloc(m.size() * sizeof(uint32_t), "link mask");
            hip_check(hipMemcpy(h->mask.p, m.data(), h->mask.bytes, hipMemcpyHostToDevice), "mask upload");
        }
        if (h->nsc) {
            upload(h->sp_cell, sf.cell.data(), sf.cell.size() * sizeof(int), "sponge cells");
            upload(h->sp_sigma, sf.sigma.data(), sf.sigma.size() * sizeof(double), "sponge strength");
            upload(h->sp_off, sf.off.data(), sf.off.size() * sizeof(int), "sponge parts");
            upload(h->sp_part, sf.part.data(), sf.part.size() * sizeof(int), "sponge parts");
            upload(h->sp_weight, sf.weight.data(), sf.weight.size() * sizeof(double), "sponge parts");
            upload(h->sp_phase, sf.phase.data(), sf.phase.size() * sizeof(double), "sponge phases");
        }
        // Collision constants.
        if (h->kind == K_MRT || h->kind == K_CUMULANT || h->kind == K_TRT_LES || h->kind == K_BGK_LES ||
            is_wale(h->kind) || is_psm(h->kind)) {
            std::vector<double> c;
            if (is_psm(h->kind)) {
                c = {h->magic, h->psm_kappa, double(h->psm_law), h->psm_c_alpha, h->psm_drag_shape};
            } else if (is_wale(h->kind)) {
                c = {h->magic, h->wale_c2, h->wale_floor};
            } else if (h->kind == K_TRT_LES || h->kind == K_BGK_LES) {
                // [magic, k = 18 sqrt(2) C^2 (psm.rs:193 evaluation order), floor]
                c = {h->magic, 18.0 * 1.4142135623730951 * h->les_c * h->les_c, h->les_floor};
            } else if (h->kind == K_MRT) {
                std::vector<double> rows, norms;
                std::vector<int> groups;
                mrt_basis(h->Q, rows, norms, groups);
                c = rows;
                c.insert(c.end(), norms.begin(), norms.end());
                for (int k = 0; k < h->Q; k++) {
                    double r = h->omega;
                    if (h->regularized) {  // bulk and shear relax with omega, the rest is removed
                        c.push_back((groups[k] == 2 || groups[k] == 3) ? h->omega : 1.0);
                        continue;
                    }
                    switch (groups[k]) {
                    case 0:
                    case 3: r = h->omega; break;
                    case 1:
                    case 4: r = h->om; break;
                    case 2: r = std::isnan(h->bulk) ? h->omega : h->bulk; break;
                    default: r = std::isnan(h->even) ? h->omega : h->even; break;
                    }
                    c.push_back(r);
                }
            } else {
                c.push_back(std::isnan(h->bulk) ? 1.0 : h->bulk);
            }
            h->h_consts = c;
            if (h->prec == LBM_F64) {
                h->mrt.alloc(c.size() * sizeof(double), "collision constants");
                hip_check(hipMemcpy(h->mrt.p, c.data(), h->mrt.bytes, hipMemcpyHostToDevice), "upload");
            } else {
                std::vector<float> cf(c.begin(), c.end());
                h->mrt.alloc(cf.size() * sizeof(float), "collision constants");
                hip_check(hipMemcpy(h->mrt.p, cf.data(), h->mrt.bytes, hipMemcpyHostToDevice), "upload");
            }
        }
        if (h->np) {
            h->pcell.alloc(h->np * sizeof(int), "ports");
            h->pnb.alloc(h->np * sizeof(int), "ports");
            h->pkind.alloc(h->np * sizeof(int), "ports");
            h->pval.alloc(h->np * 3 * sizeof(double), "ports");
            h->pbuf.alloc(size_t(h->np) * h->Q * h->csize(), "port neighbour state");
            hip_check(hipMemcpy(h->pcell.p, h->h_pcell.data(), h->pcell.bytes, hipMemcpyHostToDevice), "upload");
            hip_check(hipMemcpy(h->pnb.p, h->h_pnb.data(), h->pnb.bytes, hipMemcpyHostToDevice), "upload");
            hip_check(hipMemcpy(h->pkind.p, pkind.data(), h->pkind.bytes, hipMemcpyHostToDevice), "upload");
            hip_check(hipMemcpy(h->pval.p, pval.data(), h->pval.bytes, hipMemcpyHostToDevice), "upload");
            if (!stress_ops.empty()) {
                h->stress.alloc(stress_ops.size() * sizeof(double), "stress operators");
                hip_check(hipMemcpy(h->stress.p, stress_ops.data(), h->stress.bytes, hipMemcpyHostToDevice),
                          "upload");
            }
        }
        h->budget = cfg->checkpoint_budget_bytes;
        h->every = cfg->checkpoint_every;
        hip_check(hipStreamSynchronize(h->stream), "sync");
        *out = h.release();
        lbm_reset_initial(*out);
    });
}

int lbm_destroy(lbm_handle* h) {
    return guarded([&] {
        if (!h) return;
        if (h->stream) (void)hipStreamSynchronize(h->stream);
        if (h->ev0) (void)hipEventDestroy(h->ev0);
        if (h->ev1) (void)hipEventDestroy(h->ev1);
        hipStream_t s = h->stream;
        delete h;
        if (s) (void)hipStreamDestroy(s);
    });
}

int lbm_set_design(lbm_handle* h, const double* alpha, uint64_t n) {
    return guarded([&] {
        if (!h || !alpha) fail(LBM_ERR_CONTRACT, "null argument");
        if (n != uint64_t(h->ncell)) fail(LBM_ERR_SHAPE, "design needs one coefficient per cell");
        for (uint64_t x = 0; x < n; x++)
            if (!std::isfinite(alpha[x]) || alpha[x] < 0.0)
                fail(LBM_ERR_CONTRACT, "Brinkman coefficients must be finite and non-negative");
        h->h_alpha.assign(alpha, alpha + n);
        h->alpha_max = 0.0;
        for (uint64_t x = 0; x < n; x++)
            if (h->solid.empty() || !h->solid[x]) h->alpha_max = std::max(h->alpha_max, alpha[x]);
        if (h->prec == LBM_F64) {
            hip_check(hipMemcpy(h->alpha.p, alpha, n * sizeof(double), hipMemcpyHostToDevice), "upload");
        } else {
            std::vector<float> f(alpha, alpha + n);
            hip_check(hipMemcpy(h->alpha.p, f.data(), n * sizeof(float), hipMemcpyHostToDevice), "upload");
        }
    });
}

int lbm_set_solid_momentum(lbm_handle* h, const double* m, uint64_t len) {
    return guarded([&] {
        if (!h || !m) fail(LBM_ERR_CONTRACT, "null argument");
        if (!is_psm(h->kind)) fail(LBM_ERR_CONTRACT, "solid momentum needs the occupancy coupling");
        if (len != uint64_t(h->ncell) * 3) fail(LBM_ERR_SHAPE, "solid momentum needs three components per cell");
        const long long n = h->ncell;
        std::vector<double> soa(size_t(3 * n));
        for (long long x = 0; x < n; x++)
            for (int a = 0; a < 3; a++) {
                const double v = m[3 * x + a];
                if (!std::isfinite(v)) fail(LBM_ERR_CONTRACT, "solid momentum must be finite");
                soa[size_t(a * n + x)] = v;
            }
        if (h->prec == LBM_F64) {
            hip_check(hipMemcpy(h->sm.p, soa.data(), soa.size() * sizeof(double), hipMemcpyHostToDevice), "upload");
        } else {
            std::vector<float> f(soa.begin(), soa.end());
            hip_check(hipMemcpy(h->sm.p, f.data(), f.size() * sizeof(float), hipMemcpyHostToDevice), "upload");
        }
    });
}

int lbm_solid_momentum_bar(lbm_handle* h, double* out, uint64_t len) {
    return guarded([&] {
        if (!h || !out) fail(LBM_ERR_CONTRACT, "null argument");
        if (!is_psm(h->kind)) fail(LBM_ERR_CONTRACT, "solid momentum needs the occupancy coupling");
        if (len != uint64_t(h->ncell) * 3) fail(LBM_ERR_SHAPE, "solid momentum needs three components per cell");
        if (!h->sm_bar.p) fail(LBM_ERR_INVALID, "no adjoint sweep has run on this handle");
        const long long n = h->ncell;
        std::vector<double> soa(size_t(3 * n));
        hip_check(hipMemcpyAsync(soa.data(), h->sm_bar.p, soa.size() * sizeof(double), hipMemcpyDeviceToHost, h->stream),
                  "copy");
        hip_check(hipStreamSynchronize(h->stream), "sync");
        for (long long x = 0; x < n; x++)
            for (int a = 0; a < 3; a++) out[3 * x + a] = soa[size_t(a * n + x)];
    });
}

int lbm_set_occupancy_end(lbm_handle* h, const double* d1, const double* m1, uint64_t n, uint64_t t0,
                          uint64_t t1) {
    return guarded([&] {
        if (!h || !d1) fail(LBM_ERR_CONTRACT, "null argument");
        if (!is_psm(h->kind)) fail(LBM_ERR_CONTRACT, "an occupancy endpoint needs the occupancy coupling");
        if (n != uint64_t(h->ncell)) fail(LBM_ERR_SHAPE, "occupancy needs one value per cell");
        if (t1 <= t0) fail(LBM_ERR_CONTRACT, "the blend needs t0 < t1");
        const long long nc = h->ncell;
        std::vector<double> soa(size_t(4 * nc), 0.0);
        for (long long x = 0; x < nc; x++) {
            if (!std::isfinite(d1[x]) || d1[x] < 0.0) fail(LBM_ERR_CONTRACT, "occupancy must be finite and non-negative");
            soa[size_t(x)] = d1[x];
            for (int a = 0; a < 3; a++) {
                const double v = m1 ? m1[3 * x + a] : 0.0;
                if (!std::isfinite(v)) fail(LBM_ERR_CONTRACT, "solid momentum must be finite");
                soa[size_t((1 + a) * nc + x)] = v;
            }
        }
        h->d1.alloc(size_t(nc) * h->csize(), "second occupancy");
        h->sm1.alloc(size_t(nc) * 3 * h->csize(), "second solid momentum");
        if (h->prec == LBM_F64) {
            hip_check(hipMemcpy(h->d1.p, soa.data(), size_t(nc) * 8, hipMemcpyHostToDevice), "upload");
            hip_check(hipMemcpy(h->sm1.p, soa.data() + nc, size_t(nc) * 24, hipMemcpyHostToDevice), "upload");
        } else {
            std::vector<float> f(soa.begin(), soa.end());
            hip_check(hipMemcpy(h->d1.p, f.data(), size_t(nc) * 4, hipMemcpyHostToDevice), "upload");
            hip_check(hipMemcpy(h->sm1.p, f.data() + nc, size_t(nc) * 12, hipMemcpyHostToDevice), "upload");
        }
        h->te0 = t0;
        h->te1 = t1;
        h->sub_mid = false;
    });
}

int lbm_occupancy_end_bar(lbm_handle* h, double* d1_bar, double* m1_bar, uint64_t n) {
    return guarded([&] {
        if (!h || !d1_bar) fail(LBM_ERR_CONTRACT, "null argument");
        if (!h->d1.p) fail(LBM_ERR_CONTRACT, "no second occupancy endpoint on this handle");
        if (n != uint64_t(h->ncell)) fail(LBM_ERR_SHAPE, "occupancy needs one value per cell");
        if (!h->d1_bar.p) fail(LBM_ERR_INVALID, "no adjoint sweep has run on this handle");
        const long long nc = h->ncell;
        std::vector<double> mb(size_t(3 * nc));
        hip_check(hipMemcpyAsync(d1_bar, h->d1_bar.p, size_t(nc) * 8, hipMemcpyDeviceToHost, h->stream), "copy");
        hip_check(hipMemcpyAsync(mb.data(), h->sm1_bar.p, mb.size() * 8, hipMemcpyDeviceToHost, h->stream), "copy");
        hip_check(hipStreamSynchronize(h->stream), "sync");
        if (m1_bar)
            for (long long x = 0; x < nc; x++)
                for (int a = 0; a < 3; a++) m1_bar[3 * x + a] = mb[size_t(a * nc + x)];
    });
}

int lbm_reset_initial(lbm_handle* h) {
    return guarded([&] {
        if (!h) fail(LBM_ERR_CONTRACT, "null handle");
        // solver.rs:592-601: fluid W - 1.5 W (c.a) so the Guo velocity is 0; solids W.
        const int Q = h->Q;
        std::vector<double> fluid(Q);
        for (int q = 0; q < Q; q++) {
            double ca = 0.0;
            for (int d = 0; d < 3; d++) ca = ca + h->a[d] * double(lat_c(Q, q, d));
            fluid[q] = lat_w(Q, q) - 1.5 * lat_w(Q, q) * ca;
        }
        std::vector<double> f(size_t(h->ncell) * Q);
        for (long long x = 0; x < h->ncell; x++)
            for (int q = 0; q < Q; q++)
                f[size_t(x) * Q + q] = (!h->solid.empty() && h->solid[x]) ? lat_w(Q, q) : fluid[q];
        h->t = 0;
        upload_state(h, h->F.p, 0, f.data(), 0);
        if (h->thermal)  // T = 0 everywhere
            hip_check(hipMemsetAsync(h->tpop(h->F.p), 0, h->state_bytes - h->flow_bytes, h->stream), "memset");
        hip_check(hipStreamSynchronize(h->stream), "sync");
    });
}

int lbm_set_state(lbm_handle* h, int layout, const double* f, uint64_t len) {
    return guarded([&] {
        if (!h || !f) fail(LBM_ERR_CONTRACT, "null argument");
        if (len != uint64_t(h->ncell) * h->Q) fail(LBM_ERR_SHAPE, "state length must be cells*Q");
        upload_state(h, h->F.p, h->t, f, layout);
        hip_check(hipStreamSynchronize(h->stream), "sync");
    });
}

int lbm_get_state(lbm_handle* h, int layout, double* f, uint64_t len) {
    return guarded([&] {
        if (!h || !f) fail(LBM_ERR_CONTRACT, "null argument");
        if (len != uint64_t(h->ncell) * h->Q) fail(LBM_ERR_SHAPE, "state length must be cells*Q");
        download_state(h, h->F.p, h->t, f, layout);
    });
}

uint64_t lbm_step_count(const lbm_handle* h) { return h ? h->t : 0; }

uint64_t lbm_state_bytes(const lbm_handle* h) { return h ? h->state_bytes : 0; }

int lbm_step(lbm_handle* h, uint64_t nsteps) {
    return guarded([&] {
        if (!h) fail(LBM_ERR_CONTRACT, "null handle");
        ObjArgs none{};
        for (uint64_t s = 0; s < nsteps; s++) {
            step_once(h, h->F.p, h->t, none, nullptr, 0);
            h->t++;
        }
        hip_check(hipStreamSynchronize(h->stream), "sync");
    });
}

int lbm_forward(lbm_handle* h, uint64_t nsteps, const lbm_objective* obj, double* J, lbm_stats* stats) {
    return guarded([&] {
        if (!h) fail(LBM_ERR_CONTRACT, "null handle");
        validate_objective(h, obj, nsteps);
        const ObjArgs tmpl = objective_template(h, obj);
        const auto t0 = std::chrono::steady_clock::now();
        reset_scalars(h);
        MassLedger led;
        led.init(h, nsteps);
        hip_check(hipEventRecord(h->ev0, h->stream), "event");
        for (uint64_t s = 0; s < nsteps; s++) {
            step_once(h, h->F.p, h->t, stage_obj(tmpl, obj, s, nsteps, false), &led, s);
            h->t++;
        }
        hip_check(hipEventRecord(h->ev1, h->stream), "event");
        // Final state: objective (stage at n + terminal), its mass and diagnostics.
        StepArgs ea = final_state_args(h, h->F.p, h->t);
        ea.obj = stage_obj(tmpl, obj, nsteps, nsteps, true);
        ea.macc = led.mass(nsteps);
        ObjArgs none{};
        hip_check(h->run->objective(h->Q, h->has_mask, ea, none, nullptr, nullptr, 0, h->cfg()), "objective launch");
        thermal_final(h, h->F.p, h->t, ea.obj.tw, 0.0, 0);
        double J0 = 0.0;
        hip_check(hipMemcpyAsync(&J0, h->scal.p, sizeof(J0), hipMemcpyDeviceToHost, h->stream), "copy");
        const SweepDiag d = read_diag(h);
        const double drift = led.drift();
        if (J) *J = J0;
        if (stats) {
            std::memset(stats, 0, sizeof(*stats));
            stats->forward_steps = nsteps;
            stats->forward_ms = elapsed(h);
            stats->total_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
            stats->max_mach = d.max_mach;
            stats->bad_cells = d.bad_cells;
            stats->min_population = d.min_population;
            stats->mass_drift = drift;
            stats->forward_kernel_avg_ms = NAN;
            stats->adjoint_kernel_avg_ms = NAN;
            stats->cotangent_max = NAN;
        }
        admit(h, d, drift);
    });
}

namespace {
void check_scales(const lbm_vector_scales* sc) {
    if (!sc) fail(LBM_ERR_CONTRACT, "null scales");
    if (!(sc->velocity > 0.0) || !(sc->density > 0.0) || !(sc->step_s > 0.0) || !(sc->cell_volume > 0.0) ||
        !std::isfinite(sc->velocity + sc->density + sc->step_s + sc->cell_volume))
        fail(LBM_ERR_CONTRACT, "vector scales must be positive and finite");
}
}  // namespace

int lbm_vector(lbm_handle* h, const double* phi, uint64_t n, const lbm_vector_scales* scales, double* out) {
    return guarded([&] {
        if (!h || !out) fail(LBM_ERR_CONTRACT, "null argument");
        if (is_psm(h->kind)) fail(LBM_ERR_CONTRACT, "terminal responses read a Brinkman coefficient; not with the occupancy coupling");
        check_scales(scales);
        if (phi && n != uint64_t(h->ncell)) fail(LBM_ERR_SHAPE, "phi needs one entry per cell");
        const LaunchCfg c = h->cfg();
        const size_t nb = size_t(c.grid.x) * c.grid.y * c.grid.z;
        DevBuf sums;
        sums.alloc(nb * 5 * sizeof(double), "vector sums");
        hip_check(h->run->vector(h->Q, h->has_mask, h->g, h->F.p, h->mask.as<uint32_t>(), int(h->t & 1), h->alpha.p,
                                 h->a, scales->velocity, scales->density, 1.0 / scales->step_s, sums.as<double>(), c),
                  "vector launch");
        std::vector<double> v(nb * 5);
        hip_check(hipMemcpyAsync(v.data(), sums.p, sums.bytes, hipMemcpyDeviceToHost, h->stream), "copy");
        hip_check(hipStreamSynchronize(h->stream), "sync");
        double acc[5] = {0, 0, 0, 0, 0};
        for (size_t b = 0; b < nb; b++)
            for (int k = 0; k < 5; k++) acc[k] += v[5 * b + k];
        double vol = 0.0;
        if (phi)
            for (uint64_t x = 0; x < n; x++) vol += phi[x];
        const double fc = double(fluid_cells(h));
        out[0] = acc[0] / fc;
        out[1] = acc[1] / fc;
        out[2] = acc[2] / fc;
        out[3] = 0.5 * acc[3] * scales->cell_volume;
        out[4] = acc[4] * scales->cell_volume;
        out[5] = vol * scales->cell_volume;
    });
}

int lbm_vector_noalloc(lbm_handle* h, const double* phi_dev, uint64_t n, const lbm_vector_scales* scales,
                       double* out_dev) {
    return guarded([&] {
        if (!h || !out_dev) fail(LBM_ERR_CONTRACT, "null argument");
        if (is_psm(h->kind)) fail(LBM_ERR_CONTRACT, "terminal responses read a Brinkman coefficient; not with the occupancy coupling");
        check_scales(scales);
        if (phi_dev && n != uint64_t(h->ncell)) fail(LBM_ERR_SHAPE, "phi needs one entry per cell");
        int dev = -1;
        hip_check(hipGetDevice(&dev), "device query");
        if (dev != h->device) fail(LBM_ERR_CONTRACT, "call lbm_vector_noalloc with the handle's device current");
        const LaunchCfg c = h->cfg();
        const long long nb = (long long)c.grid.x * c.grid.y * c.grid.z;
        const size_t need = size_t(nb) * 5 + NOALLOC_PHI_BLOCKS;
        if (need > NOALLOC_SCRATCH_DOUBLES)
            fail(LBM_ERR_MEMORY, "lbm_vector_noalloc needs " + std::to_string(need) + " workspace values for this "
                                 "lattice; its fixed workspace holds " + std::to_string(NOALLOC_SCRATCH_DOUBLES) +
                                 " (use lbm_vector)");
        double* ws = scratch_address(dev);
        hip_check(h->run->vector(h->Q, h->has_mask, h->g, h->F.p, h->mask.as<uint32_t>(), int(h->t & 1), h->alpha.p,
                                 h->a, scales->velocity, scales->density, 1.0 / scales->step_s, ws, c),
                  "vector launch");
        int nphi = 0;
        if (phi_dev) {
            nphi = NOALLOC_PHI_BLOCKS;
            hipLaunchKernelGGL(k_noalloc_phi, dim3(NOALLOC_PHI_BLOCKS), dim3(256), 0, h->stream, phi_dev,
                               (long long)n, ws + size_t(nb) * 5);
            hip_check(hipGetLastError(), "volume launch");
        }
        hipLaunchKernelGGL(k_noalloc_vector, dim3(1), dim3(256), 0, h->stream, ws, nb, ws + size_t(nb) * 5, nphi,
                           1.0 / double(fluid_cells(h)), scales->cell_volume, out_dev);
        hip_check(hipGetLastError(), "fold launch");
        hip_check(hipStreamSynchronize(h->stream), "sync");
    });
}

int lbm_vector_vjp(lbm_handle* h, const double* phi, uint64_t n, const lbm_vector_scales* scales, const double* w,
                   double* f_bar, uint64_t f_len, double* alpha_bar, double* phi_bar) {
    return guarded([&] {
        if (!h || !w || !f_bar || !alpha_bar) fail(LBM_ERR_CONTRACT, "null argument");
        if (is_psm(h->kind)) fail(LBM_ERR_CONTRACT, "terminal responses read a Brinkman coefficient; not with the occupancy coupling");
        check_scales(scales);
        if (f_len != uint64_t(h->ncell) * h->Q) fail(LBM_ERR_SHAPE, "f_bar length must be cells*Q");
        if (phi && n != uint64_t(h->ncell)) fail(LBM_ERR_SHAPE, "phi needs one entry per cell");
        for (int k = 0; k < 6; k++)
            if (!std::isfinite(w[k])) fail(LBM_ERR_CONTRACT, "vector weights must be finite");
        ObjArgs ob{};
        ob.kind = OBJ_VECTOR;
        ob.weight = 1.0;
        vector_weights(h, w, *scales, ob);
        ensure_aos(h);
        DevBuf ab;
        ab.alloc(size_t(h->ncell) * sizeof(double), "vector alpha_bar");
        hip_check(h->run->vector_vjp(h->Q, h->has_mask, h->g, h->F.p, h->mask.as<uint32_t>(), int(h->t & 1),
                                     h->alpha.p, h->a, ob, h->aos.as<double>(), ab.as<double>(), h->cfg()),
                  "vector vjp launch");
        hip_check(hipMemcpyAsync(f_bar, h->aos.p, f_len * sizeof(double), hipMemcpyDeviceToHost, h->stream), "copy");
        hip_check(hipMemcpyAsync(alpha_bar, ab.p, ab.bytes, hipMemcpyDeviceToHost, h->stream), "copy");
        hip_check(hipStreamSynchronize(h->stream), "sync");
        if (phi_bar)
            for (long long x = 0; x < h->ncell; x++) phi_bar[x] = w[5] * scales->cell_volume;
        (void)phi;
    });
}

namespace {
void need_thermal(const lbm_handle* h) {
    if (!h) fail(LBM_ERR_CONTRACT, "null handle");
    if (!h->thermal) fail(LBM_ERR_CONTRACT, "no temperature field on this handle (lbm_config.thermal = 0)");
}
void ensure_thermal_aos(lbm_handle* h) {
    ensure_aos(h);  // ncell * Q doubles >= ncell * 7
}
}  // namespace

int lbm_set_thermal_design(lbm_handle* h, const double* phi, uint64_t n) {
    return guarded([&] {
        need_thermal(h);
        if (!phi) fail(LBM_ERR_CONTRACT, "null argument");
        if (n != uint64_t(h->ncell)) fail(LBM_ERR_SHAPE, "thermal design needs one value per cell");
        long long ns = 0;
        for (uint64_t x = 0; x < n; x++) {
            if (!(phi[x] >= 0.0 && phi[x] <= 1.0)) fail(LBM_ERR_CONTRACT, "thermal design values must lie in [0, 1]");
            ns += phi[x] > 0.5;
        }
        h->n_tsolid = ns;
        if (h->prec == LBM_F64) {
            hip_check(hipMemcpy(h->phi.p, phi, n * sizeof(double), hipMemcpyHostToDevice), "upload");
        } else {
            std::vector<float> f(phi, phi + n);
            hip_check(hipMemcpy(h->phi.p, f.data(), n * sizeof(float), hipMemcpyHostToDevice), "upload");
        }
    });
}

int lbm_get_thermal_state(lbm_handle* h, double* g, uint64_t len) {
    return guarded([&] {
        need_thermal(h);
        if (!g) fail(LBM_ERR_CONTRACT, "null argument");
        if (len != uint64_t(h->ncell) * TQ) fail(LBM_ERR_SHAPE, "temperature state length must be cells*7");
        ensure_thermal_aos(h);
        hip_check(h->trun->get_state(h->g, h->tpop(h->F.p), int(h->t & 1), h->aos.as<double>(), h->tdec, h->cfg()),
                  "get_state launch");
        hip_check(hipMemcpyAsync(g, h->aos.p, len * sizeof(double), hipMemcpyDeviceToHost, h->stream), "copy");
        hip_check(hipStreamSynchronize(h->stream), "sync");
    });
}

int lbm_set_thermal_state(lbm_handle* h, const double* g, uint64_t len) {
    return guarded([&] {
        need_thermal(h);
        if (!g) fail(LBM_ERR_CONTRACT, "null argument");
        if (len != uint64_t(h->ncell) * TQ) fail(LBM_ERR_SHAPE, "temperature state length must be cells*7");
        for (uint64_t x = 0; x < len; x++)
            if (!std::isfinite(g[x])) fail(LBM_ERR_CONTRACT, "temperature populations must be finite");
        ensure_thermal_aos(h);
        hip_check(hipMemcpyAsync(h->aos.p, g, len * sizeof(double), hipMemcpyHostToDevice, h->stream), "copy");
        hip_check(h->trun->set_state(h->g, h->tpop(h->F.p), int(h->t & 1), h->aos.as<double>(), h->tenc, h->cfg()),
                  "set_state launch");
        hip_check(hipStreamSynchronize(h->stream), "sync");
    });
}

int lbm_get_temperature(lbm_handle* h, double* T, uint64_t n) {
    return guarded([&] {
        need_thermal(h);
        if (!T) fail(LBM_ERR_CONTRACT, "null argument");
        if (n != uint64_t(h->ncell)) fail(LBM_ERR_SHAPE, "temperature needs one entry per cell");
        std::vector<double> g(size_t(h->ncell) * TQ);
        ensure_thermal_aos(h);
        hip_check(h->trun->get_state(h->g, h->tpop(h->F.p), int(h->t & 1), h->aos.as<double>(), h->tdec, h->cfg()),
                  "get_state launch");
        hip_check(hipMemcpyAsync(g.data(), h->aos.p, g.size() * sizeof(double), hipMemcpyDeviceToHost, h->stream),
                  "copy");
        hip_check(hipStreamSynchronize(h->stream), "sync");
        for (long long x = 0; x < h->ncell; x++) {
            double t = 0.0;
            for (int q = 0; q < TQ; q++) t += g[size_t(x) * TQ + q];
            T[x] = t;
        }
    });
}

int lbm_thermal_design_bar(lbm_handle* h, double* phi_bar, uint64_t n) {
    return guarded([&] {
        need_thermal(h);
        if (!phi_bar) fail(LBM_ERR_CONTRACT, "null argument");
        if (n != uint64_t(h->ncell)) fail(LBM_ERR_SHAPE, "phi_bar needs one entry per cell");
        if (!h->phi_bar.p) fail(LBM_ERR_INVALID, "no gradient has been computed on this handle");
        hip_check(hipMemcpy(phi_bar, h->phi_bar.p, n * sizeof(double), hipMemcpyDeviceToHost), "copy");
    });
}

int lbm_get_thermal_lambda(lbm_handle* h, double* lambda, uint64_t len) {
    return guarded([&] {
        need_thermal(h);
        if (!lambda) fail(LBM_ERR_CONTRACT, "null argument");
        if (len != uint64_t(h->ncell) * TQ) fail(LBM_ERR_SHAPE, "temperature cotangent length must be cells*7");
        if (!h->L.p) fail(LBM_ERR_INVALID, "no gradient has been computed on this handle");
        ensure_thermal_aos(h);
        hip_check(h->trun->get_lam(h->g, h->tlam(), int(h->t & 1), h->aos.as<double>(), h->cfg()), "get_lam launch");
        hip_check(hipMemcpyAsync(lambda, h->aos.p, len * sizeof(double), hipMemcpyDeviceToHost, h->stream), "copy");
        hip_check(hipStreamSynchronize(h->stream), "sync");
    });
}

int lbm_get_fields(lbm_handle* h, double* rho, double* u, uint64_t n) {
    return guarded([&] {
        if (!h || !rho || !u) fail(LBM_ERR_CONTRACT, "null argument");
        if (n != uint64_t(h->ncell)) fail(LBM_ERR_SHAPE, "fields need one entry per cell");
        DevBuf d;
        d.alloc(size_t(h->ncell) * 4 * sizeof(double), "fields");
        double* dr = d.as<double>();
        double* du = dr + h->ncell;
        hip_check(h->run->fields(h->Q, h->has_mask, h->g, h->F.p, h->mask.as<uint32_t>(), int(h->t & 1),
                                 is_psm(h->kind) ? h->alpha0.p : h->alpha.p,
                                 h->a, dr, du, h->cfg()),
                  "fields launch");
        hip_check(hipMemcpyAsync(rho, dr, n * sizeof(double), hipMemcpyDeviceToHost, h->stream), "copy");
        hip_check(hipMemcpyAsync(u, du, 3 * n * sizeof(double), hipMemcpyDeviceToHost, h->stream), "copy");
        hip_check(hipStreamSynchronize(h->stream), "sync");
    });
}

int lbm_port_cells(const lbm_handle* h, int32_t* cells, int32_t* nbs, uint64_t cap, uint64_t* count) {
    return guarded([&] {
        if (!h || !count) fail(LBM_ERR_CONTRACT, "null argument");
        *count = uint64_t(h->np);
        for (int r = 0; r < h->np && uint64_t(r) < cap; r++) {
            if (cells) cells[r] = h->h_pcell[r];
            if (nbs) nbs[r] = h->h_pnb[r];
        }
    });
}

int lbm_debug_link_table(lbm_handle* h, uint32_t* source, uint64_t len) {
    return guarded([&] {
        if (!h || !source) fail(LBM_ERR_CONTRACT, "null argument");
        const int Q = h->Q;
        if (len != uint64_t(h->ncell) * Q) fail(LBM_ERR_SHAPE, "link table length must be cells*Q");
        // Run the streaming-only operator once, in double precision, on a state whose
        // value at (x, q) is its own reference index x*Q + q: afterwards (x, q) holds its
        // source index (solver.rs:131-136).
        const Runner* r64 = runner_f64();
        DevBuf buf;
        buf.alloc(size_t(h->arr_elems) * Q * sizeof(double), "link-table state");
        hip_check(hipMemsetAsync(buf.p, 0, buf.bytes, h->stream), "memset");
        std::vector<double> ids(size_t(h->ncell) * Q);
        for (size_t i = 0; i < ids.size(); i++) ids[i] = double(i);
        ensure_aos(h);
        hip_check(hipMemcpyAsync(h->aos.p, ids.data(), ids.size() * sizeof(double), hipMemcpyHostToDevice, h->stream),
                  "copy");
        hip_check(r64->set_state(Q, h->has_mask, h->g, buf.p, h->mask.as<uint32_t>(), 0, h->aos.as<double>(), h->cfg()),
                  "set_state");
        // Built like every other step (in place on buf); the double-precision runner reads
        // a double alpha array, so give it zeros.
        StepArgs s = h->step_args(buf.p, 0);
        DevBuf zero;
        zero.alloc(size_t(h->ncell) * sizeof(double), "zero alpha");
        hip_check(hipMemsetAsync(zero.p, 0, zero.bytes, h->stream), "memset");
        s.alpha = zero.p;
        s.mrt = nullptr;
        s.tpop = nullptr;  // the flow populations only
        hip_check(r64->forward(Q, K_STREAM_ONLY, h->has_mask, s, h->cfg()), "stream-only launch");
        hip_check(r64->get_state(Q, h->has_mask, h->g, buf.p, h->mask.as<uint32_t>(), 1, h->aos.as<double>(), h->cfg()),
                  "get_state");
        hip_check(hipMemcpyAsync(ids.data(), h->aos.p, ids.size() * sizeof(double), hipMemcpyDeviceToHost, h->stream),
                  "copy");
        hip_check(hipStreamSynchronize(h->stream), "sync");
        for (long long x = 0; x < h->ncell; x++)
            for (int q = 0; q < Q; q++) {
                const size_t i = size_t(x) * Q + q;
                source[i] = (!h->solid.empty() && h->solid[x]) ? 0xFFFFFFFFu : uint32_t(ids[i]);
            }
    });
}

namespace {
void export_schedule(const Schedule& s, int64_t* actions, uint64_t cap, uint64_t* count, uint64_t* forward_steps,
                     uint64_t* reverse_start) {
    if (count) *count = s.actions.size();
    if (forward_steps) *forward_steps = s.forward_steps;
    if (reverse_start) *reverse_start = s.reverse_start;
    if (actions)
        for (size_t i = 0; i < s.actions.size() && i < cap; i++) {
            actions[5 * i + 0] = int64_t(s.actions[i].kind);
            actions[5 * i + 1] = int64_t(s.actions[i].a);
            actions[5 * i + 2] = int64_t(s.actions[i].b);
            actions[5 * i + 3] = 0;
            actions[5 * i + 4] = 0;
        }
}
}  // namespace

int lbm_every_schedule(uint64_t steps, uint64_t every, int64_t* actions, uint64_t cap, uint64_t* count,
                       uint64_t* forward_steps, uint64_t* reverse_start) {
    return guarded([&] {
        export_schedule(every_schedule(steps, every), actions, cap, count, forward_steps, reverse_start);
    });
}

int lbm_revolve_schedule(uint64_t steps, uint64_t snapshots, int64_t* actions, uint64_t cap, uint64_t* count,
                         uint64_t* forward_steps, uint64_t* reverse_start) {
    return guarded([&] {
        Schedule s = revolve_schedule(steps, snapshots);
        if (count) *count = s.actions.size();
        if (forward_steps) *forward_steps = s.forward_steps;
        if (reverse_start) *reverse_start = s.reverse_start;
        if (actions)
            for (size_t i = 0; i < s.actions.size() && i < cap; i++) {
                actions[5 * i + 0] = int64_t(s.actions[i].kind);
                actions[5 * i + 1] = int64_t(s.actions[i].a);
                actions[5 * i + 2] = int64_t(s.actions[i].b);
                actions[5 * i + 3] = 0;
                actions[5 * i + 4] = 0;
            }
    });
}

uint64_t lbm_revolve_optimal_forward_steps(uint64_t steps, uint64_t snapshots) {
    return revolve_optimal_forward_steps(steps, snapshots);
}

int lbm_reserve(lbm_handle* h, uint64_t nsteps, uint64_t* slots) {
    return guarded([&] {
        if (!h) fail(LBM_ERR_CONTRACT, "null handle");
        if (nsteps == 0) fail(LBM_ERR_INVALID, "a checkpoint schedule needs at least one step");
        const uint64_t s = reserve_pool(h, nsteps);
        if (slots) *slots = s;
    });
}

int lbm_step_vjp(lbm_handle* h, const double* lambda_next, double* lambda_prev, double* alpha_bar, uint64_t len) {
    return guarded([&] {
        if (!h || !lambda_next || !lambda_prev || !alpha_bar) fail(LBM_ERR_CONTRACT, "null argument");
        if (h->thermal)
            fail(LBM_ERR_CONTRACT, "lbm_step_vjp covers flow-only handles; use lbm_gradient with the temperature field");
        const int Q = h->Q;
        if (len != uint64_t(h->ncell) * Q) fail(LBM_ERR_SHAPE, "cotangent length must be cells*Q");
        ensure_adjoint_buffers(h);
        ensure_aos(h);
        const uint64_t t = h->t;
        // state t+1 port neighbours: step a copy of the state
        if (h->np) {
            ObjArgs none{};
            step_into(h, h->F.p, h->A.p, t, none, nullptr, 0);
            hip_check(h->run->port_save(Q, h->has_mask, h->port_args(t + 1), h->A.p, h->pbuf.p, h->stream),
                      "port save");
        }
        if (h->lam16()) {
            double mx = 0.0;
            for (uint64_t x = 0; x < len; x++) {
                if (!std::isfinite(lambda_next[x])) fail(LBM_ERR_CONTRACT, "cotangent must be finite");
                mx = std::max(mx, std::fabs(lambda_next[x]));
            }
            lam_init(h, mx, 0.0);
        }
        hip_check(hipMemcpyAsync(h->aos.p, lambda_next, len * sizeof(double), hipMemcpyHostToDevice, h->stream),
                  "copy");
        hip_check(h->run->set_lam(Q, h->has_mask, h->g, h->L.p, h->mask.as<uint32_t>(), int((t + 1) & 1),
                                  h->aos.as<double>(), h->lsc_ptr(), 0, h->cfg()),
                  "set_lam");
        if (h->np)
            hip_check(h->run->port_vjp(Q, h->has_mask, h->port_args(t + 1), h->L.p, h->pbuf.p, h->lsc_ptr(),
                                       h->stream),
                      "port vjp");
        hip_check(hipMemsetAsync(h->alpha_bar.p, 0, h->alpha_bar.bytes, h->stream), "memset");
        if (h->sm_bar.p) hip_check(hipMemsetAsync(h->sm_bar.p, 0, h->sm_bar.bytes, h->stream), "memset");
        if (h->d1_bar.p) {
            hip_check(hipMemsetAsync(h->d1_bar.p, 0, h->d1_bar.bytes, h->stream), "memset");
            hip_check(hipMemsetAsync(h->sm1_bar.p, 0, h->sm1_bar.bytes, h->stream), "memset");
        }
        AdjArgs a = h->adj_args(h->F.p, int(t & 1));
        a.theta = h->theta(t);
        flow_adjoint(h, a);
        hip_check(h->run->get_lam(Q, h->has_mask, h->g, h->L.p, h->mask.as<uint32_t>(), int(t & 1), h->aos.as<double>(),
                                  h->lsc_ptr(), h->cfg()),
                  "get_lam");
        hip_check(hipMemcpyAsync(lambda_prev, h->aos.p, len * sizeof(double), hipMemcpyDeviceToHost, h->stream),
                  "copy");
        hip_check(hipMemcpyAsync(alpha_bar, h->alpha_bar.p, size_t(h->ncell) * sizeof(double), hipMemcpyDeviceToHost,
                                 h->stream),
                  "copy");
        hip_check(hipStreamSynchronize(h->stream), "sync");
        h->abar_ready = true;
        check_lam_range(h);
    });
}

int lbm_step_jvp(lbm_handle* h, const double* d_state, const double* d_alpha, double* d_next, uint64_t len) {
    return guarded([&] {
        if (!h || !d_state || !d_next) fail(LBM_ERR_CONTRACT, "null argument");
        const int Q = h->Q;
        if (len != uint64_t(h->ncell) * Q) fail(LBM_ERR_SHAPE, "tangent length must be cells*Q");
        const bool psm_law = h->kind == K_TRT_PSM || h->kind == K_BGK_PSM;
        const bool kind_ok = h->kind == K_BGK || h->kind == K_TRT || h->kind == K_MRT || h->kind == K_BGK_GUO ||
                             h->kind == K_CUMULANT || h->kind == K_TRT_LES || h->kind == K_BGK_LES ||
                             is_wale(h->kind) || (is_psm(h->kind) && !(psm_law && h->Q == 27));
        if (!kind_ok || h->thermal)
            fail(LBM_ERR_CONTRACT, "lbm_step_jvp covers the BGK, TRT, MRT, regularized, cumulant and BGK-Guo "
                                   "Brinkman operators, the Smagorinsky and WALE closures, sponge layers and the "
                                   "occupancy couplings (the Psm law on D3Q19), without the temperature field");
        // The dual field (value, tangent) uses 16-byte elements; the twisted addressing
        // takes 32-bit byte offsets within an x plane of the field.
        if (uint64_t(h->g.SY + h->g.GY) * uint64_t(h->g.SZ) * uint64_t(Q) * sizeof(Dual) >= (uint64_t(1) << 31))
            fail(LBM_ERR_CONTRACT, "lattice too large for the tangent step (an x plane of 16-byte values above 2 GiB)");
        for (uint64_t x = 0; x < len; x++)
            if (!std::isfinite(d_state[x])) fail(LBM_ERR_CONTRACT, "tangent must be finite");
        const uint64_t t = h->t;
        const JvpRunner* jr = jvp_runner();
        const LaunchCfg lc = h->cfg();
        DevBuf dual, tan, da, dc;
        dual.alloc(size_t(h->arr_elems) * Q * sizeof(Dual), "tangent state");
        hip_check(hipMemsetAsync(dual.p, 0, dual.bytes, h->stream), "memset");
        tan.alloc(len * sizeof(double), "tangent");
        ensure_aos(h);
        // values: the state as the step reads it (decoded to FP64); tangents: the caller's
        hip_check(h->run->get_state(Q, h->has_mask, h->g, h->F.p, h->mask.as<uint32_t>(), int(t & 1),
                                    h->aos.as<double>(), lc),
                  "get_state");
        hip_check(hipMemcpyAsync(tan.p, d_state, len * sizeof(double), hipMemcpyHostToDevice, h->stream), "copy");
        hip_check(jr->set(Q, h->has_mask, h->g, dual.p, h->mask.as<uint32_t>(), int(t & 1), h->aos.as<double>(),
                          tan.as<double>(), lc.grid, lc.block, h->stream),
                  "set dual state");
        std::vector<Dual> hd(size_t(h->ncell));
        for (long long x = 0; x < h->ncell; x++) hd[x] = Dual(h->h_alpha[x], d_alpha ? d_alpha[x] : 0.0);
        da.alloc(hd.size() * sizeof(Dual), "tangent design");
        hip_check(hipMemcpyAsync(da.p, hd.data(), da.bytes, hipMemcpyHostToDevice, h->stream), "copy");
        std::vector<Dual> hc(h->h_consts.begin(), h->h_consts.end());
        if (!hc.empty()) {
            dc.alloc(hc.size() * sizeof(Dual), "tangent constants");
            hip_check(hipMemcpyAsync(dc.p, hc.data(), dc.bytes, hipMemcpyHostToDevice, h->stream), "copy");
        }
        StepArgs s = h->step_args(dual.p, t);
        s.alpha = da.p;
        s.mrt = dc.p;
        s.tpop = nullptr;
        s.wu = s.sm = s.d1 = s.sm1 = nullptr;
        DevBuf du, dsm, dd1, dsm1;
        if (is_psm(h->kind)) {  // the solid momentum (and a second endpoint) as constants
            hip_check(hipStreamSynchronize(h->stream), "sync");
            auto to_dual = [&](const DevBuf& src, size_t count, DevBuf& dst, const char* what) {
                std::vector<Dual> v(count);
                if (h->prec == LBM_F64) {
                    std::vector<double> x(count);
                    hip_check(hipMemcpy(x.data(), src.p, count * sizeof(double), hipMemcpyDeviceToHost), "copy");
                    for (size_t i = 0; i < count; i++) v[i] = Dual(x[i]);
                } else {
                    std::vector<float> x(count);
                    hip_check(hipMemcpy(x.data(), src.p, count * sizeof(float), hipMemcpyDeviceToHost), "copy");
                    for (size_t i = 0; i < count; i++) v[i] = Dual(double(x[i]));
                }
                dst.alloc(count * sizeof(Dual), what);
                hip_check(hipMemcpy(dst.p, v.data(), dst.bytes, hipMemcpyHostToDevice), "upload");
            };
            const size_t n = size_t(h->ncell);
            to_dual(h->sm, 3 * n, dsm, "tangent solid momentum");
            s.sm = dsm.p;
            if (h->d1.p) {
                to_dual(h->d1, n, dd1, "tangent second occupancy");
                to_dual(h->sm1, 3 * n, dsm1, "tangent second solid momentum");
                s.d1 = dd1.p;
                s.sm1 = dsm1.p;
            }
        }
        if (is_wale(h->kind)) {  // u and rho of every cell, with their tangents, before the in-place step
            du.alloc(size_t(h->ncell) * 4 * sizeof(Dual), "tangent WALE velocities");
            hip_check(jr->wale_u(Q, h->has_mask, h->g, dual.p, h->mask.as<uint32_t>(), int(t & 1), du.p, lc.grid,
                                 lc.block, h->stream),
                      "tangent WALE velocity launch");
            s.wu = du.p;
        }
        hip_check(jr->step(Q, h->kind, h->has_mask, s, lc.grid, lc.block, h->stream), "tangent step launch");
        if (h->nsc)  // the blend of the sponge cells' post-collision values, before the ports
            hip_check(jr->sponge(Q, h->sponge_args(t), dual.p, h->stream), "tangent sponge launch");
        {
            PortArgs pa = h->port_args(t + 1);
            pa.pre_off = 0;  // the tangent state has no pre-port region
            hip_check(jr->ports(Q, h->has_mask, pa, dual.p, h->stream), "tangent port launch");
        }
        hip_check(jr->get(Q, h->has_mask, h->g, dual.p, h->mask.as<uint32_t>(), int((t + 1) & 1), tan.as<double>(),
                          lc.grid, lc.block, h->stream),
                  "get tangent");
        hip_check(hipMemcpyAsync(d_next, tan.p, len * sizeof(double), hipMemcpyDeviceToHost, h->stream), "copy");
        hip_check(hipStreamSynchronize(h->stream), "sync");
        // solid cells keep their populations through a step (solver.rs:134): so do their tangents
        if (!h->solid.empty())
            for (long long x = 0; x < h->ncell; x++)
                if (h->solid[x])
                    for (int q = 0; q < Q; q++) d_next[size_t(x) * Q + q] = d_state[size_t(x) * Q + q];
    });
}

int lbm_alpha_bar(lbm_handle* h, double* alpha_bar, uint64_t n) {
    return guarded([&] {
        if (!h || !alpha_bar) fail(LBM_ERR_CONTRACT, "null argument");
        if (n != uint64_t(h->ncell)) fail(LBM_ERR_SHAPE, "alpha_bar needs one entry per cell");
        if (!h->abar_ready) fail(LBM_ERR_INVALID, "no gradient has been computed on this handle");
        h->stage.download(alpha_bar, h->alpha_bar.p, size_t(n) * sizeof(double), h->stream);
    });
}

int lbm_gradient(lbm_handle* h, uint64_t nsteps, const lbm_objective* obj, double* J, double* alpha_bar,
                 uint64_t n_alpha, double* lambda0, uint64_t lambda_len, lbm_stats* stats) {
    return lbm_gradient_seeded(h, nsteps, obj, nullptr, 0, J, alpha_bar, n_alpha, lambda0, lambda_len, stats);
}

int lbm_gradient_seeded(lbm_handle* h, uint64_t nsteps, const lbm_objective* obj, const double* lambda_final,
                        uint64_t lambda_final_len, double* J, double* alpha_bar, uint64_t n_alpha, double* lambda0,
                        uint64_t lambda_len, lbm_stats* stats) {
    return guarded([&] {
        if (!h) fail(LBM_ERR_CONTRACT, "null handle");
        if (nsteps == 0) fail(LBM_ERR_INVALID, "a checkpoint schedule needs at least one step");
        double lf_max = 0.0;  // max |lambda_final|
        if (lambda_final) {
            if (lambda_final_len != uint64_t(h->ncell) * h->Q)
                fail(LBM_ERR_SHAPE, "lambda_final length must be cells*Q");
            for (uint64_t x = 0; x < lambda_final_len; x++) {
                if (!std::isfinite(lambda_final[x])) fail(LBM_ERR_CONTRACT, "lambda_final must be finite");
                lf_max = std::max(lf_max, std::fabs(lambda_final[x]));
            }
            ensure_aos(h);
            hip_check(hipMemcpyAsync(h->aos.p, lambda_final, lambda_final_len * sizeof(double),
                                     hipMemcpyHostToDevice, h->stream),
                      "copy");
        }
        if (alpha_bar && n_alpha != uint64_t(h->ncell)) fail(LBM_ERR_SHAPE, "alpha_bar needs one entry per cell");
        if (lambda0 && lambda_len != uint64_t(h->ncell) * h->Q) fail(LBM_ERR_SHAPE, "lambda0 length must be cells*Q");
        validate_objective(h, obj, nsteps);
        const ObjArgs tmpl = objective_template(h, obj);
        const auto t0 = std::chrono::steady_clock::now();
        const int Q = h->Q;
        ensure_adjoint_buffers(h);
        const uint64_t base = h->t;  // state 0 of the sweep; parities follow base + s
        const uint64_t slots = reserve_pool(h, nsteps);
        Schedule sch = h->every ? every_schedule(nsteps, h->every) : revolve_schedule(nsteps, slots);
        if (sch.peak > slots)
            fail(LBM_ERR_MEMORY, "checkpoint budget holds " + std::to_string(slots) + " snapshots; checkpoint_every " +
                                     std::to_string(h->every) + " needs " + std::to_string(sch.peak));
        // State 0 stays in F for the whole sweep: its snapshot is F itself (`alias` is the
        // slot that holds it), every step writes the working buffer W, a snapshot slot or
        // the ahead buffer, and F is never written -- no copy in and no restore out.
        long long alias = -1;
        auto pool_ptr = [&](uint64_t s) { return static_cast<void*>(static_cast<char*>(h->pool.p) + size_t(s) * h->state_bytes); };
        auto slot_ptr = [&](uint64_t s) { return (long long)s == alias ? h->F.p : pool_ptr(s); };

        reset_scalars(h);
        if (h->thermal) hip_check(hipMemsetAsync(h->phi_bar.p, 0, h->phi_bar.bytes, h->stream), "memset");
        MassLedger led;
        led.init(h, nsteps);
        std::vector<hipEvent_t> evs;
        auto event = [&]() {
            hipEvent_t e;
            hip_check(hipEventCreate(&e), "event");
            hip_check(hipEventRecord(e, h->stream), "event record");
            evs.push_back(e);
            return e;
        };
        // A span that starts where the previous one ended shares its event: each record
        // leaves the stream idle for a few microseconds.  `mark` is an event at the stream's
        // current position (null once other work was queued after it).
        hipEvent_t mark = nullptr;
        auto span_start = [&]() {
            if (!mark) mark = event();
            return mark;
        };
        auto span_end = [&]() {
            mark = event();
            return mark;
        };
        double fwd_ms = 0.0, adj_ms = 0.0;
        std::vector<std::pair<hipEvent_t, hipEvent_t>> fwd_spans, adj_spans, fwd_k, adj_k;
        // Every ksample-th step kernel (forward and adjoint counted apart) is bracketed by
        // events to time the kernel alone: every 64th, and at least four of each kind on
        // short horizons (each record leaves the stream idle for a few microseconds).
        const uint64_t ksample = std::max<uint64_t>(1, std::min<uint64_t>(64, nsteps / 4));
        auto kspan = [&](std::vector<std::pair<hipEvent_t, hipEvent_t>>& v, uint64_t count)
            -> const std::pair<hipEvent_t, hipEvent_t>* {
            if (count % ksample != ksample / 2) return nullptr;
            hipEvent_t a, b;
            hip_check(hipEventCreate(&a), "event");
            hip_check(hipEventCreate(&b), "event");
            evs.push_back(a);
            evs.push_back(b);
            v.emplace_back(a, b);
            return &v.back();
        };
        uint64_t fsteps = 0, rsteps = 0, copies = 0;
        uint64_t cursor = 0;  // state index held in F
        ObjArgs none{};
        AdjArgs aa = h->adj_args(h->F.p, 0);
        uint64_t abar_run = 0;  // reverse steps gathered in the FP32 alpha_bar partial

        // The cursor state lives in `cur` (the working buffer F or a snapshot slot).  A
        // Restore only moves `cur` to the slot; the next step reads the slot and writes F
        // (out of place), and a Snapshot is taken by writing the last step of the preceding
        // advance straight into its slot -- no whole-state copies.
        void* cur = h->F.p;
        SweepDiag diag;
        double drift = 0.0;
        // The first sweep's diagnostics are read after the reverse sweep (the states they
        // cover are the first sweep's; recomputed steps record none), so the reverse sweep
        // is queued behind the forward one without a host round trip; a refused
        // trajectory fails the call all the same, with F still at state 0.
        for (size_t idx = 0; idx < sch.actions.size(); idx++) {
            const Action& act = sch.actions[idx];
            const bool forward_phase = idx < sch.reverse_start;
            switch (act.kind) {
            case ActKind::Snapshot:
                if (cursor != act.a) fail(LBM_ERR_INVALID, "snapshot of a state the cursor does not hold");
                if (act.a == 0 && cur == h->F.p && alias < 0) {  // state 0 is F
                    alias = (long long)act.b;
                    break;
                }
                if ((long long)act.b == alias) alias = -1;  // the slot now takes another state
                if (cur != slot_ptr(act.b)) {
                    hip_check(hipMemcpyAsync(slot_ptr(act.b), cur, h->state_bytes, hipMemcpyDeviceToDevice,
                                             h->stream),
                              "snapshot copy");
                    copies++;
                    mark = nullptr;
                }
                break;
            case ActKind::Restore:
                cur = slot_ptr(act.b);
                cursor = act.a;
                break;
            case ActKind::Release: break;
            case ActKind::Advance: {
                if (cursor != act.a) fail(LBM_ERR_INVALID, "advance from a position the cursor does not hold");
                // Where the advance ends: into the slot of an immediately following snapshot
                // (of a state after 0, so the slot's own storage).
                void* last_dst = h->W.p;
                if (idx + 1 < sch.actions.size() && sch.actions[idx + 1].kind == ActKind::Snapshot &&
                    sch.actions[idx + 1].a == act.b)
                    last_dst = pool_ptr(sch.actions[idx + 1].b);
                const hipEvent_t e0 = span_start();
                for (uint64_t t = act.a; t < act.b; t++) {
                    const ObjArgs ob = forward_phase ? stage_obj(tmpl, obj, t, nsteps, false) : none;
                    if (t + 1 == nsteps) {
                        // Final step into the ahead buffer; the cursor keeps state n-1.
                        step_into(h, cur, h->A.p, base + t, ob, forward_phase ? &led : nullptr, t,
                                  kspan(fwd_k, fsteps));
                        fsteps++;
                        // Objective on the final state and the cotangent seed lambda_n.
                        StepArgs ea = final_state_args(h, h->A.p, base + nsteps);
                        ea.diag = forward_phase ? 1 : 0;
                        ea.obj = forward_phase ? stage_obj(tmpl, obj, nsteps, nsteps, true) : none;
                        ea.macc = led.mass(nsteps);
                        const ObjArgs seed = stage_obj(tmpl, obj, nsteps, nsteps, true);
                        lam_init(h, seed_bound(h, seed) + lf_max,
                                 seed_bound(h, stage_obj(tmpl, obj, nsteps - 1, nsteps, false)));
                        ea.lsc = h->lsc_ptr();
                        // the seed kernel writes alpha_bar; the momentum cotangent starts at zero
                        if (h->sm_bar.p)
                            hip_check(hipMemsetAsync(h->sm_bar.p, 0, h->sm_bar.bytes, h->stream), "memset");
                        if (h->d1_bar.p) {
                            hip_check(hipMemsetAsync(h->d1_bar.p, 0, h->d1_bar.bytes, h->stream), "memset");
                            hip_check(hipMemsetAsync(h->sm1_bar.p, 0, h->sm1_bar.bytes, h->stream), "memset");
                        }
                        hip_check(h->run->objective(Q, h->has_mask, ea, seed, h->L.p, h->alpha_bar.as<double>(), 1,
                                                    h->cfg()),
                                  "seed launch");
                        thermal_final(h, h->A.p, base + nsteps, ea.obj.tw, seed.tw, 1);
                        if (lambda_final)  // lambda_n += lambda_final (the caller's terminal cotangent)
                            hip_check(h->run->set_lam(Q, h->has_mask, h->g, h->L.p, h->mask.as<uint32_t>(),
                                                      int((base + nsteps) & 1), h->aos.as<double>(), h->lsc_ptr(),
                                                      1, h->cfg()),
                                      "lambda_final upload");
                        if (h->np)
                            hip_check(h->run->port_save(Q, h->has_mask, h->port_args(base + nsteps), h->A.p,
                                                        h->pbuf.p, h->stream),
                                      "port save");
                        break;
                    }
                    void* dst = (t + 1 == act.b) ? last_dst : h->W.p;
                    step_into(h, cur, dst, base + t, ob, forward_phase ? &led : nullptr, t, kspan(fwd_k, fsteps));
                    cur = dst;
                    fsteps++;
                    cursor = t + 1;
                }
                fwd_spans.emplace_back(e0, span_end());
                break;
            }
            case ActKind::Reverse: {
                const uint64_t k = act.a;  // reverse step k-1 -> k; `cur` holds state k-1
                if (cursor + 1 != k) fail(LBM_ERR_INVALID, "reverse of a step whose states are not held");
                const hipEvent_t e0 = span_start();
                if (h->np)
                    hip_check(h->run->port_vjp(Q, h->has_mask, h->port_args(base + k), h->L.p, h->pbuf.p,
                                               h->lsc_ptr(), h->stream),
                              "port vjp");
                if (h->np && h->thermal)
                    hip_check(h->trun->port_vjp(h->port_args(base + k), h->tlam(), h->stream), "thermal port vjp");
                aa.pop = cur;
                aa.odd = int((base + k - 1) & 1);
                aa.theta = h->theta(base + k - 1);
                aa.obj = stage_obj(tmpl, obj, k - 1, nsteps, false);
                if (h->thermal) {  // lambda_g and the velocity cotangent the flow kernel adds
                    ThermalArgs ta = h->thermal_args(cur, nullptr, base + k - 1);
                    ta.obj_w = aa.obj.tw;
                    hip_check(h->trun->adjoint(Q, ta, h->cfg()), "thermal adjoint launch");
                    aa.ubar = h->ubar.p;
                }
                aa.s_next = k >= 2 ? seed_bound(h, stage_obj(tmpl, obj, k - 2, nsteps, false)) : 0.0;
                if (h->abar32.p) {  // runs of ABAR_RUN steps, the last one folded at k = 1
                    const bool fold = k == 1 || abar_run + 1 == ABAR_RUN;
                    aa.abar32 = (abar_run == 0 && fold) ? nullptr : h->abar32.as<float>();
                    aa.abar_mode = abar_run == 0 ? ABAR_FIRST : (fold ? ABAR_FOLD : ABAR_MID);
                    abar_run = fold ? 0 : abar_run + 1;
                }
                {
                    const auto* ks = kspan(adj_k, rsteps);
                    if (ks) hip_check(hipEventRecord(ks->first, h->stream), "event record");
                    flow_adjoint(h, aa, ks ? ks->second : nullptr);
                }
                rsteps++;
                if (h->np && k > 1)
                    hip_check(h->run->port_save(Q, h->has_mask, h->port_args(base + k - 1), cur, h->pbuf.p,
                                                h->stream),
                              "port save");
                adj_spans.emplace_back(e0, span_end());
                break;
            }
            }
        }
        diag = read_diag(h);
        drift = led.drift();
        try {
            admit(h, diag, drift);
        } catch (const Fail&) {  // refused: the handle is at state 0 (F), as on success
            for (auto e : evs) (void)hipEventDestroy(e);
            throw;
        }
        double J0 = 0.0;
        hip_check(hipMemcpyAsync(&J0, h->scal.p, sizeof(J0), hipMemcpyDeviceToHost, h->stream), "copy");
        if (alpha_bar) h->stage.download(alpha_bar, h->alpha_bar.p, size_t(h->ncell) * sizeof(double), h->stream);
        if (lambda0) {
            ensure_aos(h);
            hip_check(h->run->get_lam(Q, h->has_mask, h->g, h->L.p, h->mask.as<uint32_t>(), int(base & 1),
                                      h->aos.as<double>(), h->lsc_ptr(), h->cfg()),
                      "get_lam");
            h->stage.download(lambda0, h->aos.p, lambda_len * sizeof(double), h->stream);
        }
        hip_check(hipStreamSynchronize(h->stream), "sync");
        for (auto& sp : fwd_spans) {
            float ms;
            hip_check(hipEventElapsedTime(&ms, sp.first, sp.second), "event time");
            fwd_ms += ms;
        }
        for (auto& sp : adj_spans) {
            float ms;
            hip_check(hipEventElapsedTime(&ms, sp.first, sp.second), "event time");
            adj_ms += ms;
        }
        auto kavg = [&](const std::vector<std::pair<hipEvent_t, hipEvent_t>>& v) {
            double acc = 0.0;
            for (auto& sp : v) {
                float ms;
                hip_check(hipEventElapsedTime(&ms, sp.first, sp.second), "event time");
                acc += ms;
            }
            return v.empty() ? NAN : acc / double(v.size());
        };
        const double fwd_kavg = kavg(fwd_k), adj_kavg = kavg(adj_k);
        for (auto e : evs) (void)hipEventDestroy(e);
        h->t = base;
        h->abar_ready = true;
        if (J) *J = J0;
        if (stats) {
            std::memset(stats, 0, sizeof(*stats));
            stats->forward_steps = fsteps;
            stats->reverse_steps = rsteps;
            stats->snapshot_slots = slots;
            stats->peak_snapshots = sch.peak;
            stats->snapshot_copies = copies;
            stats->forward_ms = fwd_ms;
            stats->adjoint_ms = adj_ms;
            stats->total_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
            stats->max_mach = diag.max_mach;
            stats->bad_cells = diag.bad_cells;
            stats->min_population = diag.min_population;
            stats->mass_drift = drift;
            stats->forward_kernel_avg_ms = fwd_kavg;
            stats->adjoint_kernel_avg_ms = adj_kavg;
            stats->kernel_samples = fwd_k.size() + adj_k.size();
            stats->cotangent_max = NAN;
            stats->cotangent_exponent = 0;
            if (h->lam16()) {
                int rec[4];
                hip_check(hipMemcpy(rec, h->lsc.p, sizeof(rec), hipMemcpyDeviceToHost), "copy");
                float mx;
                std::memcpy(&mx, &rec[2], sizeof(mx));
                stats->cotangent_max = mx;
                stats->cotangent_exponent = rec[0];
            }
        }
        check_lam_range(h);
    });
}

// ---------------------------------------------------------------------------------
// Subcycled occupancy (SubcycledField on the moving field, moving/field.rs:1870-1919 and
// reverse/prepare :2055-2392): m fluid substeps across one macro step with the occupancy
// blended between its endpoints at theta_k = (k - 1/2) / m, the endpoints' impulse shares
// g0, g1 accumulated per cell, and the transpose by an inner binomial schedule.
// ---------------------------------------------------------------------------------
namespace {
void subcycle_check(const lbm_handle* h, uint32_t substeps) {
    if (!is_psm(h->kind)) fail(LBM_ERR_CONTRACT, "a subcycle needs the occupancy coupling");
    if (!h->d1.p) fail(LBM_ERR_CONTRACT, "a subcycle needs the second occupancy endpoint (lbm_set_occupancy_end)");
    if (substeps == 0) fail(LBM_ERR_CONTRACT, "a subcycle needs at least one substep");
    if (h->thermal) fail(LBM_ERR_CONTRACT, "a subcycle covers the flow alone");
}

// Field::check_state (moving/field.rs:1735-1775): the end state's fluid populations must be
// finite and positive and the substeps' largest speed under the Mach limit.
void admit_moving(const lbm_handle* h, const SweepDiag& steps, const SweepDiag& end) {
    if (h->mach_limit <= 0.0) return;
    if (end.bad_cells != 0) {
        char msg[200];
        std::snprintf(msg, sizeof(msg),
                      "moving LBM admission refused (populations): %llu cell states with a non-finite or "
                      "non-positive population; minimum %.3e",
                      (unsigned long long)end.bad_cells, end.min_population);
        fail(LBM_ERR_CONVERGENCE, msg);
    }
    if (steps.max_mach > h->mach_limit) {
        char msg[160];
        std::snprintf(msg, sizeof(msg), "moving LBM admission refused (mach): Mach number %.4f above the limit %g",
                      steps.max_mach, h->mach_limit);
        fail(LBM_ERR_CONVERGENCE, msg);
    }
}
}  // namespace

int lbm_subcycle(lbm_handle* h, uint32_t substeps, double* g, uint64_t g_len, lbm_stats* stats) {
    return guarded([&] {
        if (!h) fail(LBM_ERR_CONTRACT, "null handle");
        subcycle_check(h, substeps);
        const long long nc = h->ncell;
        if (g && g_len != uint64_t(nc) * 6) fail(LBM_ERR_SHAPE, "g needs 6 entries per cell");
        const auto t0 = std::chrono::steady_clock::now();
        h->gacc.grow(size_t(nc) * 6 * sizeof(double), "subcycle impulse shares");
        hip_check(hipMemsetAsync(h->gacc.p, 0, size_t(nc) * 6 * sizeof(double), h->stream), "memset");
        h->te0 = h->t;
        h->te1 = h->t + substeps;
        h->sub_mid = true;
        reset_scalars(h);
        MassLedger led;
        led.init(h, substeps);
        hip_check(hipEventRecord(h->ev0, h->stream), "event");
        h->gacc_on = true;
        try {
            for (uint32_t s = 0; s < substeps; s++) {
                step_once(h, h->F.p, h->t, ObjArgs{}, &led, s);
                h->t++;
            }
        } catch (...) {
            h->gacc_on = false;
            throw;
        }
        h->gacc_on = false;
        hip_check(hipEventRecord(h->ev1, h->stream), "event");
        const SweepDiag ds = read_diag(h);  // states 0 .. m-1: the substeps' speeds
        // the end state's populations
        hip_check(hipMemsetAsync(h->scal.as<double>() + 4, 0, 3 * sizeof(unsigned), h->stream), "memset");
        StepArgs ea = final_state_args(h, h->F.p, h->t);
        ea.macc = led.mass(substeps);
        hip_check(h->run->objective(h->Q, h->has_mask, ea, ObjArgs{}, nullptr, nullptr, 0, h->cfg()),
                  "objective launch");
        const SweepDiag de = read_diag(h);
        if (g) {
            std::vector<double> soa(size_t(nc) * 6);
            h->stage.download(soa.data(), h->gacc.p, soa.size() * sizeof(double), h->stream);
            for (long long x = 0; x < nc; x++)
                for (int e = 0; e < 6; e++) g[size_t(x) * 6 + e] = soa[size_t(e) * nc + x];
        }
        if (stats) {
            std::memset(stats, 0, sizeof(*stats));
            stats->forward_steps = substeps;
            stats->forward_ms = elapsed(h);
            stats->total_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
            stats->max_mach = ds.max_mach;
            stats->bad_cells = de.bad_cells;
            stats->min_population = de.min_population;
            stats->mass_drift = led.drift();
            stats->forward_kernel_avg_ms = NAN;
            stats->adjoint_kernel_avg_ms = NAN;
            stats->cotangent_max = NAN;
        }
        admit_moving(h, ds, de);
    });
}

int lbm_subcycle_adjoint(lbm_handle* h, uint32_t substeps, const double* state_bar, uint64_t state_len,
                         const double* g_bar, uint64_t g_len, double* previous_bar, double* d0_bar, uint64_t n,
                         lbm_stats* stats) {
    int rc = guarded([&] {
        if (!h) fail(LBM_ERR_CONTRACT, "null handle");
        subcycle_check(h, substeps);
        // The D3Q27 transpose took a GPU memory fault whose cause is not established; it
        // stays refused until it is.
        if (h->Q != 19) fail(LBM_ERR_CONTRACT, "lbm_subcycle_adjoint covers D3Q19 lattices");
        const long long nc = h->ncell;
        if (!state_bar || !previous_bar) fail(LBM_ERR_CONTRACT, "null argument");
        if (state_len != uint64_t(nc) * h->Q) fail(LBM_ERR_SHAPE, "state_bar length must be cells*Q");
        if (g_bar && g_len != uint64_t(nc) * 6) fail(LBM_ERR_SHAPE, "g_bar needs 6 entries per cell");
        if (d0_bar && n != uint64_t(nc)) fail(LBM_ERR_SHAPE, "d0_bar needs one entry per cell");
        h->te0 = h->t;
        h->te1 = h->t + substeps;
        h->sub_mid = true;
        h->gbar_on = false;
        if (g_bar) {
            std::vector<double> soa(size_t(nc) * 6);
            for (long long x = 0; x < nc; x++)
                for (int e = 0; e < 6; e++) {
                    const double v = g_bar[size_t(x) * 6 + e];
                    if (!std::isfinite(v)) fail(LBM_ERR_CONTRACT, "g_bar must be finite");
                    soa[size_t(e) * nc + x] = v;
                }
            h->gbar.grow(soa.size() * sizeof(double), "subcycle impulse cotangents");
            hip_check(hipMemcpy(h->gbar.p, soa.data(), soa.size() * sizeof(double), hipMemcpyHostToDevice),
                      "upload");
            h->gbar_on = true;
        }
    });
    if (rc != LBM_OK) return rc;
    // The transpose of the m substeps seeded with state_bar, through the gradient driver
    // (its schedule is the inner one: binomial within the checkpoint budget).
    rc = lbm_gradient_seeded(h, substeps, nullptr, state_bar, state_len, nullptr, d0_bar, n, previous_bar, state_len,
                             stats);
    h->gbar_on = false;
    return rc;
}

int lbm_time_kernel(lbm_handle* h, int which, uint64_t launches, double* avg_ms) {
    return guarded([&] {
        if (!h || !avg_ms) fail(LBM_ERR_CONTRACT, "null argument");
        if (launches == 0) fail(LBM_ERR_CONTRACT, "launches must be positive");
        if (which < 0 || which > 7) fail(LBM_ERR_CONTRACT, "unknown kernel selector");
        if (which == 4) {
            // thermal adjoint kernel alone, in place (the state is not meaningful afterwards)
            if (!h->thermal) fail(LBM_ERR_CONTRACT, "no temperature field on this handle");
            {
                ensure_adjoint_buffers(h);
                hip_check(hipMemsetAsync(h->L.p, 0, h->L.bytes, h->stream), "memset");
                hip_check(hipMemsetAsync(h->alpha_bar.p, 0, h->alpha_bar.bytes, h->stream), "memset");
                if (h->sm_bar.p) hip_check(hipMemsetAsync(h->sm_bar.p, 0, h->sm_bar.bytes, h->stream), "memset");
                if (h->d1_bar.p) {
                    hip_check(hipMemsetAsync(h->d1_bar.p, 0, h->d1_bar.bytes, h->stream), "memset");
                    hip_check(hipMemsetAsync(h->sm1_bar.p, 0, h->sm1_bar.bytes, h->stream), "memset");
                }
                hip_check(hipMemsetAsync(h->phi_bar.p, 0, h->phi_bar.bytes, h->stream), "memset");
            }
            auto launch = [&](uint64_t s) {
                ThermalArgs ta = h->thermal_args(h->F.p, h->F.p, h->t + s);
                hip_check(h->trun->adjoint(h->Q, ta, h->cfg()), "thermal launch");
            };
            launch(0);  // untimed
            hip_check(hipEventRecord(h->ev0, h->stream), "event");
            for (uint64_t s = 0; s < launches; s++) launch(s + 1);
            hip_check(hipEventRecord(h->ev1, h->stream), "event");
        } else if (which == 5 || which == 6) {
            // the sponge blend pass (5, in place on the live state, which it moves towards
            // the far field) or its transpose (6, on a zero cotangent) alone
            if (!h->nsc) fail(LBM_ERR_CONTRACT, "no sponge layer on this handle");
            if (which == 6) {
                ensure_adjoint_buffers(h);
                hip_check(hipMemsetAsync(h->L.p, 0, h->L.bytes, h->stream), "memset");
                lam_init(h, 0.0, 0.0);
            }
            auto launch = [&](uint64_t s) {
                const SpongeArgs sa = h->sponge_args(h->t + s);
                if (which == 5)
                    hip_check(h->run->sponge(h->Q, sa, h->F.p, nullptr, h->stream), "sponge launch");
                else
                    hip_check(h->run->sponge_vjp(h->Q, sa, h->L.p, h->lsc_ptr(), h->stream), "sponge transpose launch");
            };
            launch(0);  // untimed
            hip_check(hipEventRecord(h->ev0, h->stream), "event");
            for (uint64_t s = 0; s < launches; s++) launch(s + 1);
            hip_check(hipEventRecord(h->ev1, h->stream), "event");
        } else if (which == 0 || which == 2 || which == 3 || which == 7) {
            // forward steps in place on the live state (no ports: the kernel alone); which
            // 2 runs the same addressing with the collision replaced by a copy (the
            // streaming-only memory ceiling of the layout); which 3 the flow kernel that
            // also advances the temperature field (0 and 2 leave it out)
            if (which == 3 && !h->thermal) fail(LBM_ERR_CONTRACT, "no temperature field on this handle");
            const int kind = which == 2 ? int(K_STREAM_ONLY) : h->kind;
            auto args = [&](uint64_t t) {
                StepArgs sa = h->step_args(h->F.p, t);
                if (which != 3) sa.tpop = nullptr;
                if (which == 7) {  // a first-sweep step: admission diagnostics, partials folded
                    sa.diag = 1;
                    sa.part = h->part.as<double>();
                }
                return sa;
            };
            if (is_wale(h->kind))  // velocities the WALE kernel reads (the pre-pass itself is not timed)
                hip_check(h->run->wale_u(h->Q, h->has_mask, h->g, h->F.p, h->mask.as<uint32_t>(), int(h->t & 1),
                                         h->wu.p, h->cfg()),
                          "WALE velocity launch");
            {  // untimed first launch (code object load)
                StepArgs sa = args(h->t);
                hip_check(h->run->forward(h->Q, kind, h->has_mask, sa, h->cfg()), "forward launch");
                h->t++;
            }
            hip_check(hipEventRecord(h->ev0, h->stream), "event");
            for (uint64_t s = 0; s < launches; s++) {
                StepArgs sa = args(h->t);
                hip_check(h->run->forward(h->Q, kind, h->has_mask, sa, h->cfg()), "forward launch");
                h->t++;
            }
            hip_check(hipEventRecord(h->ev1, h->stream), "event");
        } else {
            ensure_adjoint_buffers(h);
            hip_check(hipMemsetAsync(h->L.p, 0, h->L.bytes, h->stream), "memset");
            hip_check(hipMemsetAsync(h->alpha_bar.p, 0, h->alpha_bar.bytes, h->stream), "memset");
            if (h->sm_bar.p) hip_check(hipMemsetAsync(h->sm_bar.p, 0, h->sm_bar.bytes, h->stream), "memset");
            if (h->d1_bar.p) {
                hip_check(hipMemsetAsync(h->d1_bar.p, 0, h->d1_bar.bytes, h->stream), "memset");
                hip_check(hipMemsetAsync(h->sm1_bar.p, 0, h->sm1_bar.bytes, h->stream), "memset");
            }
            lam_init(h, 0.0, 0.0);
            AdjArgs a = h->adj_args(h->F.p, 0);
            a.odd = int(h->t & 1);
            // FP32 arithmetic: the timed launches are the middle of a gradient's run of
            // reverse steps (alpha_bar gathered in the FP32 partial)
            a.abar32 = h->abar32.as<float>();
            a.abar_mode = ABAR_FIRST;
            flow_adjoint(h, a);  // untimed
            a.abar_mode = ABAR_MID;
            hip_check(hipEventRecord(h->ev0, h->stream), "event");
            for (uint64_t s = 0; s < launches; s++) {
                a.odd = int((h->t + s) & 1);
                flow_adjoint(h, a);
            }
            hip_check(hipEventRecord(h->ev1, h->stream), "event");
        }
        hip_check(hipStreamSynchronize(h->stream), "sync");
        *avg_ms = double(elapsed(h)) / double(launches);
    });
}

}  // extern "C"

// Plain copy kernel used as the bandwidth reference: one 16-byte element per thread.
__global__ __launch_bounds__(256) void k_copy4(const float4* __restrict__ a, float4* __restrict__ b, size_t n) {
    const size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) b[i] = a[i];
}

extern "C" int lbm_copy_bandwidth(uint64_t bytes, uint64_t iters, double* gbps, double* avg_ms) {
    return guarded([&] {
        if (!gbps || iters == 0 || bytes < 16) fail(LBM_ERR_CONTRACT, "invalid copy benchmark arguments");
        const size_t n = bytes / 16;
        DevBuf a, b;
        a.alloc(n * 16, "copy source");
        b.alloc(n * 16, "copy destination");
        hip_check(hipMemset(a.p, 0, a.bytes), "memset");
        hipEvent_t e0, e1;
        hip_check(hipEventCreate(&e0), "event");
        hip_check(hipEventCreate(&e1), "event");
        if (n / 256 + 1 > size_t(0x7fffffff)) fail(LBM_ERR_CONTRACT, "copy benchmark size too large");
        const unsigned blocks = unsigned((n + 255) / 256);
        hipLaunchKernelGGL(k_copy4, dim3(blocks), dim3(256), 0, 0, a.as<float4>(), b.as<float4>(), n);
        hip_check(hipGetLastError(), "copy launch");
        hip_check(hipEventRecord(e0, 0), "event");
        for (uint64_t i = 0; i < iters; i++)
            hipLaunchKernelGGL(k_copy4, dim3(blocks), dim3(256), 0, 0, a.as<float4>(), b.as<float4>(), n);
        hip_check(hipGetLastError(), "copy launch");
        hip_check(hipEventRecord(e1, 0), "event");
        hip_check(hipEventSynchronize(e1), "sync");
        float ms = 0.0f;
        hip_check(hipEventElapsedTime(&ms, e0, e1), "event time");
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        const double avg = double(ms) / double(iters);
        *gbps = 2.0 * double(n) * 16.0 / (avg * 1e-3) / 1e9;
        if (avg_ms) *avg_ms = avg;
    });
}

// ---------------------------------------------------------------------------------
// Design map (design.rs:95-280) on the device.
// ---------------------------------------------------------------------------------
struct lbm_design_map {
    int nx = 0, ny = 0, nz = 0;
    int per[3] = {0, 0, 0};
    long long n = 0;
    int noff = 0;
    double beta = 0.0, eta = 0.5, pa = 0.0, pden = 1.0;
    DevBuf off, w, flags, fixed, raw, phi, bar, scale;
    hipStream_t stream = nullptr;
    ~lbm_design_map() {
        if (stream) (void)hipStreamDestroy(stream);
    }
    DesignMapArgs args() const {
        DesignMapArgs a{};
        a.nx = nx;
        a.ny = ny;
        a.nz = nz;
        a.px = per[0];
        a.py = per[1];
        a.pz = per[2];
        a.noff = noff;
        a.off = off.as<int>();
        a.w = w.as<double>();
        a.flags = flags.as<uint8_t>();
        a.fixed = fixed.as<double>();
        a.beta = beta;
        a.eta = eta;
        a.pa = pa;
        a.pden = pden;
        return a;
    }
    LaunchCfg cfg() const { return LaunchCfg{dim3(unsigned((n + 255) / 256)), dim3(256), stream}; }
};

extern "C" {

int lbm_design_map_create(const lbm_design_map_config* cfg, const uint8_t* solid, const uint8_t* region,
                          const double* fixed, uint64_t n, lbm_design_map** out) {
    return guarded([&] {
        if (!cfg || !out) fail(LBM_ERR_CONTRACT, "null config or output");
        *out = nullptr;
        if (cfg->struct_size != sizeof(lbm_design_map_config))
            fail(LBM_ERR_CONTRACT, "lbm_design_map_config struct_size mismatch");
        auto m = std::make_unique<lbm_design_map>();
        for (int d = 0; d < 3; d++) {
            if (cfg->shape[d] < 1 || cfg->shape[d] > 65535) fail(LBM_ERR_CONTRACT, "design map shape out of range");
            m->per[d] = cfg->periodic[d] ? 1 : 0;
        }
        m->nx = cfg->shape[0];
        m->ny = cfg->shape[1];
        m->nz = cfg->shape[2];
        m->n = (long long)m->nx * m->ny * m->nz;
        if (n != uint64_t(m->n)) fail(LBM_ERR_SHAPE, "design map masks need one entry per cell");
        // TopologyMap::normalize (design.rs:51-80) with the radius in cells
        if (!(cfg->radius >= 0.0 && cfg->radius <= 4.0)) fail(LBM_ERR_CONTRACT, "filter radius must lie in [0,4] cells");
        if (!(cfg->beta >= 0.0 && cfg->beta <= 32.0) || !(cfg->eta > 0.0 && cfg->eta < 1.0))
            fail(LBM_ERR_CONTRACT, "Require beta in [0,32] and eta in (0,1)");
        m->beta = cfg->beta;
        m->eta = cfg->eta;
        if (m->beta > 0.0) {
            m->pa = std::tanh(m->beta * m->eta);
            m->pden = m->pa + std::tanh(m->beta * (1.0 - m->eta));
        }
        // the cone's offsets and weights in the reference's loop order (design.rs:117-133)
        std::vector<int> off;
        std::vector<double> w;
        if (cfg->radius > 0.0) {
            const long long reach = (long long)std::ceil(cfg->radius);
            for (long long a = -reach; a <= reach; a++)
                for (long long b = -reach; b <= reach; b++)
                    for (long long c = -reach; c <= reach; c++) {
                        const double norm = std::sqrt(double(a * a + b * b + c * c));
                        const double weight = std::max(cfg->radius - norm, 0.0);
                        if (weight != 0.0) {
                            off.push_back(int(a));
                            off.push_back(int(b));
                            off.push_back(int(c));
                            w.push_back(weight);
                        }
                    }
        }
        m->noff = int(w.size());
        std::vector<uint8_t> fl(size_t(m->n), 0);
        for (long long x = 0; x < m->n; x++) {
            const bool sol = solid && solid[x];
            const bool reg = region ? region[x] != 0 : true;
            fl[size_t(x)] = uint8_t((sol ? DM_SOLID : 0) | (reg && !sol ? DM_EDIT : 0));
            if (fixed && !std::isfinite(fixed[x])) fail(LBM_ERR_CONTRACT, "fixed design values must be finite");
        }
        hip_check(hipStreamCreateWithFlags(&m->stream, hipStreamNonBlocking), "stream create");
        if (m->noff) {
            upload(m->off, off.data(), off.size() * sizeof(int), "design map offsets");
            upload(m->w, w.data(), w.size() * sizeof(double), "design map weights");
        }
        upload(m->flags, fl.data(), fl.size(), "design map flags");
        m->fixed.alloc(size_t(m->n) * sizeof(double), "fixed design");
        if (fixed)
            hip_check(hipMemcpy(m->fixed.p, fixed, m->fixed.bytes, hipMemcpyHostToDevice), "upload");
        else
            hip_check(hipMemset(m->fixed.p, 0, m->fixed.bytes), "memset");
        m->raw.alloc(size_t(m->n) * sizeof(double), "raw design");
        m->phi.alloc(size_t(m->n) * sizeof(double), "mapped design");
        *out = m.release();
    });
}

int lbm_design_map_destroy(lbm_design_map* m) {
    delete m;
    return LBM_OK;
}

int lbm_design_map_apply(lbm_design_map* m, const double* raw, double* phi, uint64_t n) {
    return guarded([&] {
        if (!m || !raw || !phi) fail(LBM_ERR_CONTRACT, "null argument");
        if (n != uint64_t(m->n)) fail(LBM_ERR_SHAPE, "design needs one value per cell");
        const size_t bytes = size_t(m->n) * sizeof(double);
        hip_check(hipMemcpyAsync(m->raw.p, raw, bytes, hipMemcpyHostToDevice, m->stream), "copy");
        const LaunchCfg c = m->cfg();
        hipLaunchKernelGGL(k_design_map, c.grid, c.block, 0, c.stream, m->args(), m->raw.as<double>(),
                           m->phi.as<double>(), (long long)m->n);
        hip_check(hipGetLastError(), "design map launch");
        hip_check(hipMemcpyAsync(phi, m->phi.p, bytes, hipMemcpyDeviceToHost, m->stream), "copy");
        hip_check(hipStreamSynchronize(m->stream), "sync");
    });
}

int lbm_design_map_vjp(lbm_design_map* m, const double* raw, const double* phi_bar, double* raw_bar, uint64_t n) {
    return guarded([&] {
        if (!m || !raw || !phi_bar || !raw_bar) fail(LBM_ERR_CONTRACT, "null argument");
        if (n != uint64_t(m->n)) fail(LBM_ERR_SHAPE, "design needs one value per cell");
        const size_t bytes = size_t(m->n) * sizeof(double);
        m->bar.alloc(bytes, "design cotangent");
        m->scale.alloc(bytes, "design map scale");
        hip_check(hipMemcpyAsync(m->raw.p, raw, bytes, hipMemcpyHostToDevice, m->stream), "copy");
        hip_check(hipMemcpyAsync(m->bar.p, phi_bar, bytes, hipMemcpyHostToDevice, m->stream), "copy");
        const LaunchCfg c = m->cfg();
        hipLaunchKernelGGL(k_design_scale, c.grid, c.block, 0, c.stream, m->args(), m->raw.as<double>(),
                           m->bar.as<double>(), m->scale.as<double>(), (long long)m->n);
        hip_check(hipGetLastError(), "design scale launch");
        hipLaunchKernelGGL(k_design_gather, c.grid, c.block, 0, c.stream, m->args(), m->scale.as<double>(),
                           m->bar.as<double>(), (long long)m->n);
        hip_check(hipGetLastError(), "design gather launch");
        hip_check(hipMemcpyAsync(raw_bar, m->bar.p, bytes, hipMemcpyDeviceToHost, m->stream), "copy");
        hip_check(hipStreamSynchronize(m->stream), "sync");
    });
}

}  // extern "C"

// ---------------------------------------------------------------------------------
// Push-forward of body points (moving/pushforward.rs) on the device.
// ---------------------------------------------------------------------------------
namespace {

PushArgs push_args(const lbm_pushforward_config* cfg, const double* xi, const double* a, const double* v,
                   uint64_t np, uint64_t ncells) {
    if (!cfg) fail(LBM_ERR_CONTRACT, "null config");
    if (cfg->struct_size != sizeof(lbm_pushforward_config))
        fail(LBM_ERR_CONTRACT, "lbm_pushforward_config struct_size mismatch");
    PushArgs p{};
    int n[3];
    for (int d = 0; d < 3; d++) {
        if (cfg->shape[d] < 1 || cfg->shape[d] > 65535) fail(LBM_ERR_CONTRACT, "push-forward shape out of range");
        n[d] = cfg->shape[d];
    }
    p.nx = n[0];
    p.ny = n[1];
    p.nz = n[2];
    p.px = cfg->periodic[0] ? 1 : 0;
    p.py = cfg->periodic[1] ? 1 : 0;
    p.pz = cfg->periodic[2] ? 1 : 0;
    if (ncells != uint64_t(n[0]) * n[1] * n[2]) fail(LBM_ERR_SHAPE, "occupancy needs one entry per cell");
    if (cfg->width_cells < 1 || cfg->width_cells > uint32_t(PF_MAX_H))
        fail(LBM_ERR_CONTRACT, "kernel half-width must be 1 or 2 cells");
    if (cfg->kernel > 1) fail(LBM_ERR_CONTRACT, "unknown interpolation kernel");
    // Frame::new and Frame::with_mirrors (pushforward.rs:113-165)
    const int span = 4 * int(cfg->width_cells);
    for (int e = 0; e < 3; e++) {
        const bool pe = cfg->periodic[e] != 0;
        if (cfg->mirrored[e] > 3) fail(LBM_ERR_CONTRACT, "mirror flags are bits 0 and 1");
        if (pe && n[e] > 1 && n[e] < span)
            fail(LBM_ERR_CONTRACT, "periodic axis " + std::to_string(e) + " (" + std::to_string(n[e]) +
                                       " cells) is shorter than the push-forward kernel support (" +
                                       std::to_string(span) + " cells)");
        if (!pe && n[e] == 1)
            fail(LBM_ERR_CONTRACT,
                 "a one-cell lattice axis (" + std::to_string(e) + ") must be periodic (planar) for the push-forward");
        if (cfg->mirrored[e] && (pe || n[e] < span))
            fail(LBM_ERR_CONTRACT, "push-forward mirror face on axis " + std::to_string(e) +
                                       " needs a non-periodic axis of at least " + std::to_string(span) + " cells");
        p.mirror[e] = cfg->mirrored[e];
    }
    p.h = int(cfg->width_cells);
    p.peskin = int(cfg->kernel);
    p.np = (long long)np;
    if (np && (!xi || !a)) fail(LBM_ERR_CONTRACT, "null point arrays");
    const int per[3] = {p.px, p.py, p.pz};
    const double hf = double(p.h);
    for (uint64_t q = 0; q < np; q++) {
        if (!std::isfinite(a[q])) fail(LBM_ERR_CONTRACT, "point weights must be finite");
        for (int e = 0; e < 3; e++) {
            const double x = xi[3 * q + e];
            if (!std::isfinite(x)) fail(LBM_ERR_CONTRACT, "push-forward point " + std::to_string(q) + " is not finite");
            if (v && !std::isfinite(v[3 * q + e])) fail(LBM_ERR_CONTRACT, "point velocities must be finite");
            if (per[e] && std::fabs(x) >= 1.0e9)
                fail(LBM_ERR_CONTRACT, "push-forward point " + std::to_string(q) + " lies too far from the lattice");
            // Stencils::new_with_kernel (pushforward.rs:209-236)
            if (per[e]) continue;
            const bool lo = p.mirror[e] & 1, hi = p.mirror[e] & 2;
            if ((lo && x < -0.5) || (hi && x > double(n[e]) - 0.5))
                fail(LBM_ERR_CONTRACT, "moving LBM admission refused (pushforward_support): point " + std::to_string(q) +
                                           " on axis " + std::to_string(e) + " lies beyond a symmetry face");
            if ((!lo && x - 2.0 * hf <= -1.0) || (!hi && x + 2.0 * hf >= double(n[e])))
                fail(LBM_ERR_CONTRACT, "moving LBM admission refused (pushforward_support): point " + std::to_string(q) +
                                           " has kernel support outside the lattice on axis " + std::to_string(e));
        }
    }
    return p;
}

}  // namespace

extern "C" {

int lbm_pushforward(const lbm_pushforward_config* cfg, const double* xi, const double* a, const double* v,
                    uint64_t npoints, double* d, double* m, uint64_t ncells) {
    return guarded([&] {
        if (!d) fail(LBM_ERR_CONTRACT, "null occupancy output");
        PushArgs p = push_args(cfg, xi, a, v, npoints, ncells);
        DevBuf bxi, ba, bv, bd, bm;
        bd.alloc(size_t(ncells) * sizeof(double), "occupancy");
        hip_check(hipMemset(bd.p, 0, bd.bytes), "memset");
        if (m) {
            bm.alloc(size_t(ncells) * 3 * sizeof(double), "solid momentum");
            hip_check(hipMemset(bm.p, 0, bm.bytes), "memset");
        }
        if (npoints) {
            if (npoints >= (uint64_t(1) << 31) || ncells >= (uint64_t(1) << 31))
                fail(LBM_ERR_CONTRACT, "push-forward needs fewer than 2^31 points and cells");
            upload(bxi, xi, size_t(npoints) * 3 * sizeof(double), "points");
            upload(ba, a, size_t(npoints) * sizeof(double), "point weights");
            if (v && m) upload(bv, v, size_t(npoints) * 3 * sizeof(double), "point velocities");
            p.xi = bxi.as<double>();
            p.a = ba.as<double>();
            p.v = bv.as<double>();
            const int s = 4 * p.h;
            DevBuf key, idx, skey, sidx, wts, vel, start, end, temp;
            key.alloc(size_t(npoints) * sizeof(int), "point keys");
            idx.alloc(size_t(npoints) * sizeof(int), "point order");
            skey.alloc(size_t(npoints) * sizeof(int), "sorted keys");
            sidx.alloc(size_t(npoints) * sizeof(int), "sorted order");
            wts.alloc(size_t(npoints) * 3 * s * sizeof(double), "sorted point weights");
            if (m) vel.alloc(size_t(npoints) * 3 * sizeof(double), "sorted point velocities");
            start.alloc(size_t(ncells) * sizeof(int), "bin starts");
            end.alloc(size_t(ncells) * sizeof(int), "bin ends");
            hip_check(hipMemset(start.p, 0, start.bytes), "memset");
            hip_check(hipMemset(end.p, 0, end.bytes), "memset");
            const unsigned pblocks = unsigned((npoints + 255) / 256);
            hipLaunchKernelGGL(k_pf_keys, dim3(pblocks), dim3(256), 0, 0, p, key.as<int>(), idx.as<int>());
            hip_check(hipGetLastError(), "push-forward keys launch");
            int bits = 1;
            while ((uint64_t(1) << bits) < ncells) bits++;
            size_t temp_bytes = 0;
            hip_check(hipcub::DeviceRadixSort::SortPairs(nullptr, temp_bytes, key.as<int>(), skey.as<int>(),
                                                         idx.as<int>(), sidx.as<int>(), int(npoints), 0, bits),
                      "radix sort size");
            temp.alloc(std::max<size_t>(temp_bytes, 16), "radix sort scratch");
            hip_check(hipcub::DeviceRadixSort::SortPairs(temp.p, temp_bytes, key.as<int>(), skey.as<int>(),
                                                         idx.as<int>(), sidx.as<int>(), int(npoints), 0, bits),
                      "radix sort");
            hipLaunchKernelGGL(k_pf_bins, dim3(pblocks), dim3(256), 0, 0, p, skey.as<int>(), sidx.as<int>(),
                               start.as<int>(), end.as<int>(), wts.as<double>(), m ? vel.as<double>() : nullptr);
            hip_check(hipGetLastError(), "push-forward bins launch");
            hipLaunchKernelGGL(k_pf_gather, dim3(unsigned((ncells + 255) / 256)), dim3(256), 0, 0, p, start.as<int>(),
                               end.as<int>(), wts.as<double>(), m ? vel.as<double>() : nullptr, bd.as<double>(),
                               m ? bm.as<double>() : nullptr, (long long)ncells);
            hip_check(hipGetLastError(), "push-forward gather launch");
        }
        hip_check(hipMemcpy(d, bd.p, bd.bytes, hipMemcpyDeviceToHost), "copy");
        if (m) hip_check(hipMemcpy(m, bm.p, bm.bytes, hipMemcpyDeviceToHost), "copy");
    });
}

int lbm_pushforward_vjp(const lbm_pushforward_config* cfg, const double* xi, const double* a, const double* v,
                        uint64_t npoints, const double* d_bar, const double* m_bar, uint64_t ncells, double* xi_bar,
                        double* a_bar, double* v_bar) {
    return guarded([&] {
        if (npoints && (!xi_bar || !a_bar)) fail(LBM_ERR_CONTRACT, "null cotangent output");
        PushArgs p = push_args(cfg, xi, a, v, npoints, ncells);
        if (npoints == 0) return;
        DevBuf bxi, ba, bv, bdb, bmb, oxi, oa, ov;
        upload(bxi, xi, size_t(npoints) * 3 * sizeof(double), "points");
        upload(ba, a, size_t(npoints) * sizeof(double), "point weights");
        if (v) upload(bv, v, size_t(npoints) * 3 * sizeof(double), "point velocities");
        if (d_bar) upload(bdb, d_bar, size_t(ncells) * sizeof(double), "occupancy cotangent");
        if (m_bar) upload(bmb, m_bar, size_t(ncells) * 3 * sizeof(double), "momentum cotangent");
        oxi.alloc(size_t(npoints) * 3 * sizeof(double), "point cotangents");
        oa.alloc(size_t(npoints) * sizeof(double), "weight cotangents");
        if (v_bar) ov.alloc(size_t(npoints) * 3 * sizeof(double), "velocity cotangents");
        p.xi = bxi.as<double>();
        p.a = ba.as<double>();
        p.v = bv.as<double>();
        hipLaunchKernelGGL(k_pushforward_vjp, dim3(unsigned((npoints + 127) / 128)), dim3(128), 0, 0, p,
                           bdb.as<double>(), bmb.as<double>(), oxi.as<double>(), oa.as<double>(), ov.as<double>());
        hip_check(hipGetLastError(), "push-forward transpose launch");
        hip_check(hipMemcpy(xi_bar, oxi.p, oxi.bytes, hipMemcpyDeviceToHost), "copy");
        hip_check(hipMemcpy(a_bar, oa.p, oa.bytes, hipMemcpyDeviceToHost), "copy");
        if (v_bar) hip_check(hipMemcpy(v_bar, ov.p, ov.bytes, hipMemcpyDeviceToHost), "copy");
    });
}

}  // extern "C"
