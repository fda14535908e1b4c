// nbh_api.hip -- host side of the engine and the C ABI declared in include/nbh.h.
//
// Plan building follows build_fp_plan (src/fpntt.jl:236-271): the same stage
// list, roots, Winograd constants (src/fpntt.jl:127-168), 1/N and wi, computed
// in host FP64 with the same operation sequence; the twiddle tables are built
// on the device (k_twiddle_stage).  The schedule -- which consecutive stages
// share one LDS tile -- is this engine's own and is set by the tuning knobs.
// Sizing restates ntt_len/fp_ntt_params (src/fpntt.jl:933-977).

#include <hip/hip_runtime.h>

#include <dlfcn.h>

#include <algorithm>
#include <cstdio>
#include <cstring>
#include <deque>
#include <functional>
#include <map>
#include <memory>
#include <mutex>
#include <new>
#include <stdexcept>
#include <string>
#include <thread>
#include <tuple>
#include <vector>

#include "nbh.h"
#include "nbh_field.h"
#include "nbh_internal.h"
#include "nbh_launch.h"

using namespace nbh;

// A fixed operand's finished forward spectra (both primes) at one length and
// chunk width.
// Workspace of nbh_mul_dev_noalloc, for callers that cannot allocate device
// memory: a static device array, so it exists on every GPU the library runs
// on until the process ends (each GPU holds its own copy).  Host code never
// uses its name as an address: nbh_init looks up each device's address once
// (Lib::scratch_dev) and every launch gets that.  Its contents persist from
// call to call; nothing in it needs clearing (the pack writes every spectrum
// point and every unpack chunk writes its carry summary before it is read).
constexpr size_t NOALLOC_SCRATCH_BYTES = (size_t)64 << 20;
static __device__ __align__(256) unsigned char scratch[NOALLOC_SCRATCH_BYTES];
constexpr int SCRATCH_MAX_DEVICES = 64;

struct nbh_spectrum_s {
    const uint64_t *b = nullptr;
    int64_t n = 0, bits_b = 0, bits_a_max = 0, N = 0, ncb = 0;
    int32_t bch = 0, zero = 0;
    double *x[2] = {nullptr, nullptr};
};

namespace {

typedef unsigned __int128 u128;

thread_local std::string t_err;

int fail(int code, const std::string &msg) {
    t_err = msg;
    return code;
}

#define HIPCHK(expr)                                                                        \
    do {                                                                                    \
        hipError_t e_ = (expr);                                                             \
        if (e_ != hipSuccess)                                                               \
            return fail(e_ == hipErrorOutOfMemory ? NBH_ERR_OOM : NBH_ERR_HIP,              \
                        std::string(#expr) + ": " + hipGetErrorString(e_));                 \
    } while (0)

// Entry points return an error code for every failure: no C++ exception
// (a failed host allocation) crosses the C ABI.
#define NBH_GUARD_BEGIN try {
#define NBH_GUARD_END                                                          \
    }                                                                          \
    catch (const std::bad_alloc &) {                                           \
        return fail(NBH_ERR_OOM, "host memory allocation failed");             \
    }                                                                          \
    catch (const std::exception &ex_) {                                        \
        return fail(NBH_ERR_ARG, std::string("unexpected failure: ") + ex_.what()); \
    }                                                                          \
    catch (...) {                                                              \
        return fail(NBH_ERR_ARG, "unexpected failure");                        \
    }

#define RCCHK(expr)                \
    do {                           \
        int rc_ = (expr);          \
        if (rc_) return rc_;       \
    } while (0)

uint64_t powmod(uint64_t b, uint64_t e, uint64_t m) {
    u128 r = 1, x = b % m;
    while (e) {
        if (e & 1) r = r * x % m;
        x = x * x % m;
        e >>= 1;
    }
    return (uint64_t)r;
}

uint64_t invmod(uint64_t a, uint64_t m) {
    __int128 t = 0, nt = 1, r = m, nr = a % m;
    while (nr) {
        __int128 q = r / nr, tmp;
        tmp = t - q * nt; t = nt; nt = tmp;
        tmp = r - q * nr; r = nr; nr = tmp;
    }
    if (t < 0) t += m;
    return (uint64_t)t;
}

int top_set_bit(uint64_t x) { return x ? 64 - __builtin_clzll(x) : 0; }
int top_set_bit128(u128 x) {
    uint64_t hi = (uint64_t)(x >> 64);
    return hi ? 64 + top_set_bit(hi) : top_set_bit((uint64_t)x);
}

uint64_t prime_u(int k) { return k ? P2_U : P1_U; }
double prime_d(int k) { return k ? P2_D : P1_D; }
int two_adicity(int k) { return k ? 38 : 34; }

// fp_generator (src/fpntt.jl:56-62): smallest g in 3..63 of full order.
uint64_t generator(int k) {
    static const uint64_t f1[] = {2, 3, 5, 7, 23}, f2[] = {2, 3, 5, 7, 13};
    const uint64_t *f = k ? f2 : f1;
    const uint64_t p = prime_u(k);
    for (uint64_t g = 3; g <= 63; g++) {
        bool ok = true;
        for (int i = 0; i < 5; i++)
            if (powmod(g, (p - 1) / f[i], p) == 1) { ok = false; break; }
        if (ok) return g;
    }
    return 0;
}

// ntt_len (src/fpntt.jl:933-951); -1 where the reference throws ArgumentError.
int64_t ntt_len(int64_t T) {
    static const int64_t s[] = {1, 3, 5, 9};
    static const int64_t md[] = {1, 3, 5, 7, 9, 15, 21, 105};
    static const int64_t lg[] = {1, 3, 5, 7, 9, 15, 21, 35, 45, 63, 105, 315};
    const int64_t *ms = T <= (1ll << 14) ? s : T <= (1ll << 20) ? md : lg;
    const int nms = T <= (1ll << 14) ? 4 : T <= (1ll << 20) ? 8 : 12;
    const int maxk = std::min(two_adicity(0), two_adicity(1));
    int64_t best = INT64_MAX;
    for (int i = 0; i < nms; i++) {
        const int64_t m = ms[i];
        const int64_t cl = (T + m - 1) / m;
        int k = top_set_bit((uint64_t)(cl - 1));
        if (k < 2) k = 2;
        if (k > maxk) continue;
        best = std::min(best, m << k);
    }
    return best == INT64_MAX ? -1 : best;
}

// fp_ntt_params (src/fpntt.jl:964-977).
int ntt_params(int64_t bits_a, int64_t bits_b, int *b_out, int64_t *N_out) {
    const u128 pp = (u128)P1_U * P2_U - 1;
    const int maxb = top_set_bit(std::min(P1_U, P2_U)) - 1;
    const int64_t mn = std::min(bits_a, bits_b);
    const int nclow = top_set_bit((uint64_t)((mn + maxb - 1) / maxb));
    const int bstart = std::min(maxb, (top_set_bit128(pp) - nclow + 3) >> 1);
    for (int b = bstart; b >= 1; b--) {
        const int64_t nca = (bits_a + b - 1) / b, ncb = (bits_b + b - 1) / b;
        const u128 d = (u128)((1ull << b) - 1) * ((1ull << b) - 1);
        if ((u128)std::min(nca, ncb) <= pp / d) {
            const int64_t N = ntt_len(nca + ncb - 1);
            if (N < 0) return fail(NBH_ERR_LENGTH, "fp NTT length exceeds the 315*2^34 the primes support");
            *b_out = b;
            *N_out = N;
            return NBH_OK;
        }
    }
    return fail(NBH_ERR_LENGTH, "no admissible chunk width");
}

// Winograd constants (src/fpntt.jl:127-168), host FP64, same op sequence.
void dft5_consts(uint64_t c, int k, double *out) {
    const uint64_t p = prime_u(k);
    const double inv2 = (double)invmod(2, p);
    const double c1 = fp_reduce_k((double)c, k);
    const double c2 = fp_mulmod3_k(c1, c1, k), c3 = fp_mulmod3_k(c2, c1, k), c4 = fp_mulmod3_k(c3, c1, k);
    const double inv4 = fp_mulmod3_k(inv2, inv2, k);
    out[0] = fp_minbal_k(-inv4, k);
    out[1] = fp_minbal_k(fp_mulmod3_k((c1 + c4) - (c2 + c3), inv4, k), k);
    out[2] = fp_minbal_k(fp_mulmod3_k(c1 - c4, inv2, k), k);
    out[3] = fp_minbal_k(fp_mulmod3_k(c2 - c3, inv2, k), k);
}

void dft7_consts(uint64_t c, int k, double *out) {
    const uint64_t p = prime_u(k);
    const double inv2 = (double)invmod(2, p), inv3 = (double)invmod(3, p);
    double cj[7];
    for (int j = 1; j <= 6; j++) cj[j] = (double)powmod(c, (uint64_t)j, p);
    auto A = [&](int j) { return fp_mulmod3_k(cj[j] + cj[7 - j], inv2, k); };
    auto B = [&](int j) { return fp_mulmod3_k(cj[j] - cj[7 - j], inv2, k); };
    const double a0 = A(1), a1 = A(3), a2 = A(2);
    const double b0 = B(1), b1 = B(3), b2 = B(2);
    const double bp0 = fp_minbal_k(fp_mulmod3_k(a0 - a2, inv3, k), k);
    const double bp1 = fp_minbal_k(fp_mulmod3_k(a1 - a2, inv3, k), k);
    const double n0 = fp_minbal_k(fp_mulmod3_k(b0 - b2, inv3, k), k);
    const double n1 = fp_minbal_k(fp_mulmod3_k(b1 + b2, inv3, k), k);
    out[0] = fp_minbal_k(fp_mulmod3_k((a0 + a1) + a2, inv3, k), k);
    out[1] = bp0;
    out[2] = bp1;
    out[3] = fp_minbal_k(bp0 + bp1, k);
    out[4] = fp_minbal_k(fp_mulmod3_k((b0 + b2) - b1, inv3, k), k);
    out[5] = n0;
    out[6] = n1;
    out[7] = fp_minbal_k(n0 + n1, k);
}

struct Tuning {
    int grp_lds = 16384;  // points per strided group tile
    int mid_max = 8192;   // largest contiguous last group (one LDS array in the mid kernel)
    int cmin = 8;
    int cmax = 64;
};

// Stage list of one (N, prime), its stages pointing at the shared twiddle
// tables (plus the contiguous copy for short lengths); schedule-free.
struct StageList {
    int64_t N = 0, N2 = 0;
    int prime = 0;
    std::vector<StageDesc> st;
    double ninv = 0, ninvp = 0;
    double *twc = nullptr;  // contiguous copy of the stages' rows (N <= SMALL_TWC_MAX_N)
    int64_t ntw = 0;        // doubles of all the stages' rows
};

struct PlanEntry {
    DevPlan host;
    DevPlan *dev = nullptr;
    size_t grp_lds_bytes[NBH_MAX_GROUPS] = {};
    int grp_threads[NBH_MAX_GROUPS] = {};
    size_t mid_lds_bytes[2] = {};  // [mul (two operands), sqr (one)]
    int mid_threads[2] = {};       // [mul, sqr]
};

struct Lib {
    std::recursive_mutex mu;
    bool init = false;
    int device = -1;
    hipStream_t stream = nullptr;
    hipEvent_t ev_in = nullptr, ev_out = nullptr;
    // end of the last call that used the shared single-product resources
    // (workspace, I/O buffers), on whatever stream it ran
    hipEvent_t last = nullptr;
    bool last_valid = false;
    bool last_on_lib = false;  // `last` was recorded on g.stream
    int cus = 0;
    size_t lds_block = 65536;
    size_t hbm = 0;
    std::string name;
    Tuning tun;                                    // default schedule knobs
    std::vector<std::pair<int64_t, Tuning>> table;  // per-length knobs: first entry with N <= Nmax
    // knobs for products run alone (nbh_mul & co.): their blocks must spread one
    // transform over the chip, so they favour smaller tiles than batches do
    std::vector<std::pair<int64_t, Tuning>> table_single;
    std::map<std::pair<int64_t, int>, std::unique_ptr<StageList>> stages;
    // shared twiddle tables by (prime, radix, span)
    std::map<std::tuple<int, int, int64_t>, double *> twtab;
    std::map<std::tuple<int64_t, int, int, int, int, int>, std::unique_ptr<PlanEntry>> plans;
    void *ws = nullptr;  // single-product workspace
    size_t ws_size = 0;
    void *io = nullptr;  // host-pointer API staging of a, b, r
    size_t io_size = 0;
    void *pin = nullptr;  // pinned host bounce buffer of the host-pointer API
    void *pin_dev = nullptr;  // its device address (null: kernels do not read it directly)
    size_t pin_size = 0;
    // host-pointer batches: two staging slots (pinned operands + results, and
    // the device copy with the batch's arena), so one group's host copies
    // overlap the device work of the one before
    struct HostSlot {
        void *pin = nullptr, *dev = nullptr;
        size_t pin_size = 0, dev_size = 0;
        hipEvent_t ev = nullptr;
    } hs[2];
    hipStream_t side[8] = {};  // batch phases fork onto these (default priority)
    int nside = 8;             // how many of them a step uses (0: all launches in order)
    // per-kernel event timing
    bool prof = false;
    std::vector<std::tuple<int, hipEvent_t, hipEvent_t>> prof_open;
    std::vector<hipEvent_t> prof_pool;
    std::vector<hipEvent_t> ev_pool;  // fork/join events of released batches
    // handles given to callers and not yet destroyed (batches, cofactor steps,
    // spectra, Barrett contexts and batches, servers): they point into the plan
    // cache and twiddle tables, so nbh_shutdown refuses while any is live
    int64_t live = 0;
    // bumped whenever a schedule knob changes: arena sizes depend on it
    uint64_t tuning_gen = 0;
    // device address of `scratch` per device ordinal (looked up at nbh_init)
    void *scratch_dev[SCRATCH_MAX_DEVICES] = {};
};
Lib g;

hipEvent_t prof_event() {
    hipEvent_t e = nullptr;
    if (!g.prof_pool.empty()) {
        e = g.prof_pool.back();
        g.prof_pool.pop_back();
    } else if (hipEventCreate(&e) != hipSuccess) {
        e = nullptr;
    }
    return e;
}

// Bracket one launch with events on its stream when profiling is on.
struct ProfScope {
    int kind;
    hipStream_t s;
    hipEvent_t a = nullptr;
    ProfScope(int k, hipStream_t st, bool on = true) : kind(k), s(st) {
        if (on && g.prof && (a = prof_event())) (void)hipEventRecord(a, s);
    }
    ~ProfScope() {
        if (!a) return;
        hipEvent_t b = prof_event();
        if (!b) {
            g.prof_pool.push_back(a);
            return;
        }
        (void)hipEventRecord(b, s);
        g.prof_open.emplace_back(kind, a, b);
    }
};
// Side streams a batch step may fork onto.  HIP maps streams onto its
// hardware queues (4 by default); eight side streams over four queues beat
// four (batch 3.99 -> 3.94 ms A/B), while eight hardware queues made every
// setting slower (4.2 ms).
constexpr int NSIDE = 8;

int ensure_init() {
    if (g.init) return NBH_OK;
    return fail(NBH_ERR_NOT_INIT, "nbh_init has not been called");
}

// The shared twiddle table of a radix-m stage of span L for prime k: rows
// t = 1..m-1 of minbal(w_L^(t j)), j < L/m, with w_L = g^((p-1)/L) (the
// reference's w_N^(N/L) for any N the stage occurs in, src/fpntt.jl:173-225).
// Built on the device from exact lo/hi power tables of w_L.
int get_twtab(int k, int m, int64_t L, const double **out) {
    const auto key = std::make_tuple(k, m, L);
    auto it = g.twtab.find(key);
    if (it != g.twtab.end()) {
        *out = it->second;
        return NBH_OK;
    }
    const uint64_t p = prime_u(k);
    const uint64_t w = powmod(generator(k), (p - 1) / (uint64_t)L, p);
    const int64_t Q = L / m;
    const int logL = top_set_bit((uint64_t)(L - 1));
    const int h = (logL + 1) / 2;
    const int64_t nlo = 1ll << h;
    const int64_t nhi = (L + nlo - 1) / nlo;
    std::vector<double> tab((size_t)(nlo + nhi));
    const double half = prime_d(k) / 2;
    auto minbal = [&](uint64_t v) { return (double)v > half ? (double)v - prime_d(k) : (double)v; };
    uint64_t x = 1;
    for (int64_t i = 0; i < nlo; i++) {
        tab[(size_t)i] = minbal(x);
        x = (uint64_t)((u128)x * w % p);
    }
    const uint64_t step = x;  // w^(2^h)
    x = 1;
    for (int64_t i = 0; i < nhi; i++) {
        tab[(size_t)(nlo + i)] = minbal(x);
        x = (uint64_t)((u128)x * step % p);
    }
    double *t = nullptr, *dtab = nullptr;
    HIPCHK(hipMalloc((void **)&t, sizeof(double) * (size_t)((m - 1) * Q)));
    g.twtab[key] = t;
    HIPCHK(hipMalloc((void **)&dtab, sizeof(double) * tab.size()));
    HIPCHK(hipMemcpyAsync(dtab, tab.data(), sizeof(double) * tab.size(), hipMemcpyHostToDevice, g.stream));
    HIPCHK(launch_twiddle_stage(k, t, m, Q, 1, dtab, dtab + nlo, h, g.stream));
    HIPCHK(hipStreamSynchronize(g.stream));
    HIPCHK(hipFree(dtab));
    *out = t;
    return NBH_OK;
}

// Every stage's shared table, and (short lengths) the contiguous copy of the
// rows the one-workgroup kernel loads in one round.
int build_twiddles(StageList &L) {
    int64_t total = 0;
    for (auto &s : L.st) {
        s.tw = nullptr;
        s.tw_off = 0;
        if (s.m == 2) continue;
        RCCHK(get_twtab(L.prime, s.m, s.L, &s.tw));
        s.tw_off = total;
        total += (int64_t)(s.m - 1) * s.Q;
    }
    L.ntw = total;
    if (total == 0 || L.N > SMALL_TWC_MAX_N) return NBH_OK;
    HIPCHK(hipMalloc((void **)&L.twc, sizeof(double) * (size_t)total));
    for (auto &s : L.st)
        if (s.m != 2)
            HIPCHK(hipMemcpyAsync(L.twc + s.tw_off, s.tw, sizeof(double) * (size_t)((s.m - 1) * s.Q),
                                  hipMemcpyDeviceToDevice, g.stream));
    HIPCHK(hipStreamSynchronize(g.stream));
    return NBH_OK;
}

// build_fp_plan (src/fpntt.jl:236-271), plus the trailing radix-2 pass that
// fp_ntt_fwd!/rev! apply when k is odd (src/fpntt.jl:715-733) as a stage.
int get_stages(int64_t N, int k, StageList **out) {
    auto key = std::make_pair(N, k);
    auto it = g.stages.find(key);
    if (it != g.stages.end()) {
        *out = it->second.get();
        return NBH_OK;
    }
    const uint64_t p = prime_u(k);
    if (N < 4) return fail(NBH_ERR_LENGTH, "transform length below 4");
    const int kk = __builtin_ctzll((uint64_t)N);
    int64_t m = N >> kk;
    const bool m_ok = m == 1 || m == 3 || m == 5 || m == 7 || m == 9 || m == 15 || m == 21 || m == 35 || m == 45 ||
                      m == 63 || m == 105 || m == 315;
    if (!m_ok || !(m == 1 || kk >= 2) || kk > two_adicity(k) || (p - 1) % (uint64_t)N != 0)
        return fail(NBH_ERR_LENGTH, "unsupported transform length " + std::to_string(N));
    auto L = std::make_unique<StageList>();
    L->N = N;
    L->prime = k;
    const uint64_t w = powmod(generator(k), (p - 1) / (uint64_t)N, p);
    int64_t span = N;
    uint64_t r = w;
    const int mfs[4] = {7, 5, 3, 3};
    for (int mf : mfs) {
        if (m % mf) continue;
        StageDesc s{};
        s.m = mf;
        s.L = span;
        s.Q = span / mf;
        const uint64_t c = powmod(r, (uint64_t)s.Q, p);
        if (mf == 5) dft5_consts(c, k, s.rot);
        else if (mf == 7) dft7_consts(c, k, s.rot);
        else
            for (int t = 1; t < mf; t++) s.rot[t - 1] = fp_reduce_k((double)powmod(c, (uint64_t)t, p), k);
        L->st.push_back(s);
        r = powmod(r, (uint64_t)mf, p);
        span /= mf;
        m /= mf;
    }
    L->N2 = span;
    const double wi = span >= 4 ? fp_reduce_k((double)powmod(r, (uint64_t)(span >> 2), p), k) : 1.0;
    for (int64_t Lspan = span; Lspan >= 4; Lspan >>= 2) {
        StageDesc s{};
        s.m = 4;
        s.L = Lspan;
        s.Q = Lspan / 4;
        s.rot[0] = wi;
        L->st.push_back(s);
    }
    if (__builtin_ctzll((uint64_t)span) & 1) {
        StageDesc s{};
        s.m = 2;
        s.L = 2;
        s.Q = 1;
        L->st.push_back(s);
    }
    if ((int)L->st.size() > NBH_MAX_STAGES) return fail(NBH_ERR_LENGTH, "too many stages");
    L->ninv = fp_reduce_k((double)invmod((uint64_t)N % p, p), k);
    L->ninvp = L->ninv / prime_d(k);
    RCCHK(build_twiddles(*L));
    *out = L.get();
    g.stages[key] = std::move(L);
    return NBH_OK;
}

int64_t pow2_part(int64_t x) { return x & -x; }
int64_t floor_pow2(int64_t x) { return x < 1 ? 0 : (int64_t)1 << (63 - __builtin_clzll((uint64_t)x)); }

// The column-kernel shape of a strided group (radices in forward order as
// M0 | M1 << 4 | M2 << 8), or 0 when the column kernels do not take it: one
// to three stages of radix 3, 4, 5 or 7 with at most 36 rows (the shapes
// col_switch lists).
int col_code(const StageList &L, const GroupDesc &G) {
    if (G.S > 36 || G.s1 - G.s0 > 3) return 0;
    int code = 0;
    for (int si = G.s0; si < G.s1; si++) {
        const int m = L.st[si].m;
        if (m != 3 && m != 4 && m != 5 && m != 7) return 0;
        code |= m << (4 * (si - G.s0));
    }
    switch (code) {
    case 0x3: case 0x4: case 0x5: case 0x7: case 0x33: case 0x43: case 0x35: case 0x44:
    case 0x45: case 0x37: case 0x47: case 0x57: case 0x433: return code;
    default: return 0;
    }
}

// Cut the stage list into HBM passes: the last group is every stage of span
// <= mid_max (contiguous blocks); the earlier stages are packed greedily into
// strided groups whose S x C tile fits grp_lds doubles with C >= cmin columns.
int make_schedule(const StageList &L, const Tuning &T, DevPlan &P, PlanEntry &E) {
    const int nst = (int)L.st.size();
    int sl = 0;
    while (sl < nst && L.st[sl].L > T.mid_max) sl++;
    if (sl == nst) return fail(NBH_ERR_ARG, "schedule: no stage fits the mid budget");
    int ng = 0;
    auto colw = [&](int64_t Lb, int64_t S) {
        return std::max<int64_t>(1, std::min({pow2_part(Lb), floor_pow2(T.grp_lds / S), (int64_t)T.cmax}));
    };
    int s = 0;
    while (s < sl) {
        int64_t S = L.st[s].m;
        int e = s + 1;
        while (e < sl) {
            const int64_t S2 = S * L.st[e].m;
            if (S2 > T.grp_lds || colw(L.st[e].Q, S2) < T.cmin) break;
            S = S2;
            e++;
        }
        if (ng >= NBH_MAX_GROUPS - 1) return fail(NBH_ERR_ARG, "schedule: too many groups");
        GroupDesc &G = P.g[ng++];
        G.s0 = s;
        G.s1 = e;
        G.La = L.st[s].L;
        G.Lb = L.st[e - 1].Q;
        G.S = (int32_t)S;
        G.C = (int32_t)colw(G.Lb, S);
        G.logC = __builtin_ctz((unsigned)G.C);
        G.tiles = (L.N / G.La) * (G.Lb / G.C);
        s = e;
    }
    GroupDesc &G = P.g[ng++];
    G.s0 = sl;
    G.s1 = nst;
    G.La = L.st[sl].L;
    G.Lb = 1;
    G.S = (int32_t)G.La;
    G.C = 1;
    G.logC = 0;
    G.tiles = L.N / G.La;
    P.ngroups = ng;
    for (int gi = 0; gi < ng; gi++) {
        GroupDesc &Gd = P.g[gi];
        for (int si = Gd.s0; si < Gd.s1; si++) {
            StageDesc &sd = P.st[si];
            sd.q = (int32_t)(sd.Q / Gd.Lb);
            sd.q_magic = ((1ull << 32) + (uint64_t)sd.q - 1) / (uint64_t)sd.q;
        }
        E.grp_lds_bytes[gi] = sizeof(double) * (size_t)lds_doubles(Gd.S * Gd.C);
        E.grp_threads[gi] = tile_threads((int64_t)Gd.S * Gd.C);
        // a strided group of at most 36 rows in a supported shape runs as a
        // column group: one thread per column, up to COL_THREADS columns a tile
        const int code = gi + 1 < ng ? col_code(L, Gd) : 0;
        if (code) {
            Gd.col = code;
            Gd.C = (int32_t)std::min<int64_t>(pow2_part(Gd.Lb), COL_THREADS);
            Gd.logC = __builtin_ctz((unsigned)Gd.C);
            Gd.tiles = (L.N / Gd.La) * (Gd.Lb / Gd.C);
            E.grp_lds_bytes[gi] = Gd.S <= 16 ? COL_SMALL : COL_BIG;
            E.grp_threads[gi] = Gd.C;
        }
    }
    // the mid kernel holds one array per tile (products keep b's last pass in registers)
    E.mid_lds_bytes[0] = sizeof(double) * (size_t)lds_doubles(P.g[ng - 1].S);
    E.mid_lds_bytes[1] = sizeof(double) * (size_t)lds_doubles(P.g[ng - 1].S);
    // one thread per 16-point sub-problem of ONE array: the mid kernel runs a
    // product's two operands in the same thread
    E.mid_threads[0] = tile_threads((int64_t)P.g[ng - 1].S);
    E.mid_threads[1] = tile_threads((int64_t)P.g[ng - 1].S);
    if (E.mid_threads[0] > MID_MAX_THREADS) return fail(NBH_ERR_ARG, "schedule: mid tile too large");
    return NBH_OK;
}

Tuning tuning_for(int64_t N, bool single = false) {
    if (single)
        for (const auto &e : g.table_single)
            if (N <= e.first) return e.second;
    for (const auto &e : g.table)
        if (N <= e.first) return e.second;
    return g.tun;
}

bool tuning_ok(const Tuning &t) {
    auto pow2 = [](int32_t x) { return x > 0 && (x & (x - 1)) == 0; };
    return pow2(t.grp_lds) && pow2(t.mid_max) && pow2(t.cmin) && pow2(t.cmax) && t.cmin <= t.cmax &&
           t.mid_max >= 4 && t.grp_lds >= 16 && t.mid_max <= (1 << 16) && t.grp_lds <= (1 << 16) &&
           8 * (size_t)lds_doubles(t.grp_lds) <= g.lds_block && 8 * (size_t)lds_doubles(t.mid_max) <= g.lds_block &&
           t.mid_max <= 16 * MID_MAX_THREADS;
}

// The autotuner's table, "Nmax grp_lds mid_max cmin cmax" per line, read at
// init from tuning_<arch>.txt beside this library (if present).
void load_tuning_file() {
    Dl_info di;
    if (!dladdr((void *)&load_tuning_file, &di) || !di.dli_fname) return;
    std::string dir(di.dli_fname);
    const size_t slash = dir.rfind('/');
    dir = slash == std::string::npos ? "." : dir.substr(0, slash);
    std::string arch = g.name.substr(0, g.name.find(':'));
    FILE *f = std::fopen((dir + "/tuning_" + arch + ".txt").c_str(), "r");
    if (!f) return;
    // "Nmax grp_lds mid_max cmin cmax" for batches; the same after "single"
    // for products run alone
    std::vector<std::pair<int64_t, Tuning>> tab, tab1;
    char line[256];
    bool ok = true;
    while (std::fgets(line, sizeof line, f)) {
        if (line[0] == '#' || line[0] == '\n') continue;
        const bool single = std::strncmp(line, "single ", 7) == 0;
        auto &t_ = single ? tab1 : tab;
        long long nmax;
        Tuning t;
        if (std::sscanf(line + (single ? 7 : 0), "%lld %d %d %d %d", &nmax, &t.grp_lds, &t.mid_max, &t.cmin,
                        &t.cmax) != 5 ||
            !tuning_ok(t) || (!t_.empty() && nmax <= t_.back().first)) {
            ok = false;
            break;
        }
        t_.emplace_back((int64_t)nmax, t);
    }
    std::fclose(f);
    if (ok) {
        g.table = tab;
        g.table_single = tab1;
    }
}

int get_plan(int64_t N, int k, PlanEntry **out, bool single = false) {
    const Tuning T = tuning_for(N, single);
    auto key = std::make_tuple(N, k, T.grp_lds, T.mid_max, T.cmin, T.cmax);
    auto it = g.plans.find(key);
    if (it != g.plans.end()) {
        *out = it->second.get();
        return NBH_OK;
    }
    StageList *L = nullptr;
    RCCHK(get_stages(N, k, &L));
    auto E = std::make_unique<PlanEntry>();
    DevPlan &P = E->host;
    std::memset(&P, 0, sizeof P);
    P.N = N;
    P.N2 = L->N2;
    P.prime = k;
    P.nst = (int32_t)L->st.size();
    P.ninv = L->ninv;
    P.ninvp = L->ninvp;
    for (int i = 0; i < P.nst; i++) P.st[i] = L->st[i];
    P.twc = L->twc;
    RCCHK(make_schedule(*L, T, P, *E));
    HIPCHK(hipMalloc((void **)&E->dev, sizeof(DevPlan)));
    HIPCHK(hipMemcpy(E->dev, &P, sizeof(DevPlan), hipMemcpyHostToDevice));
    *out = E.get();
    g.plans[key] = std::move(E);
    return NBH_OK;
}

// The cached plan of (N, k) under the current schedule, or null: never builds
// one (building allocates the plan record and the twiddle tables).
PlanEntry *find_plan(int64_t N, int k, bool single) {
    const Tuning T = tuning_for(N, single);
    auto it = g.plans.find(std::make_tuple(N, k, T.grp_lds, T.mid_max, T.cmin, T.cmax));
    return it == g.plans.end() ? nullptr : it->second.get();
}

size_t align256(size_t x) { return (x + 255) & ~(size_t)255; }

// A caller-independent workspace for products that must not allocate: the
// library's static device array (its device address and size).
struct Scratch {
    void *p;
    size_t bytes;
};

// Host-pointer products at most this large (operands + result) go through
// the pinned bounce buffer: a host memcpy plus a DMA from pinned memory beats
// a copy from pageable memory at these sizes; larger ones copy directly.
constexpr size_t PIN_MAX_BYTES = (size_t)64 << 20;
// From this size on (a 1024 x 1024-limb product), a host-pointer product's
// kernels read the operands from and write the result to the pinned buffer
// directly: two copy commands cost more than the bus latency the kernels then
// see (4096 x 4096 limbs: 58 us per call instead of 78; below 1024 limbs the
// copies are 1-2 us faster).
constexpr size_t ZERO_COPY_MIN_BYTES = (size_t)32 << 10;

int ensure_pinned(size_t need) {
    if (g.pin_size >= need) return NBH_OK;
    HIPCHK(hipDeviceSynchronize());
    if (g.pin) HIPCHK(hipHostFree(g.pin));
    g.pin = g.pin_dev = nullptr;
    g.pin_size = 0;
    const size_t sz = std::max(need, (size_t)1 << 20);
    HIPCHK(hipHostMalloc(&g.pin, sz, hipHostMallocDefault));
    g.pin_size = sz;
    if (hipHostGetDevicePointer(&g.pin_dev, g.pin, 0) != hipSuccess) {
        (void)hipGetLastError();
        g.pin_dev = nullptr;
    }
    return NBH_OK;
}

int ensure_buf(void **buf, size_t *cur, size_t need) {
    if (*cur >= need) return NBH_OK;
    HIPCHK(hipDeviceSynchronize());
    if (*buf) HIPCHK(hipFree(*buf));
    *buf = nullptr;
    *cur = 0;
    const size_t sz = std::max(need, (size_t)1 << 20);
    HIPCHK(hipMalloc(buf, sz));
    *cur = sz;
    return NBH_OK;
}

int64_t bits_from_top(int64_t n, uint64_t top) { return 64 * (n - 1) + top_set_bit(top); }

// Fill a DevJob's sizing fields; returns NBH_OK with J.zero set for a zero operand.
int size_job(DevJob &J, int64_t bits_a, int64_t bits_b, int64_t *N_out) {
    J.rn = J.m + J.n;
    J.nchunks = (J.rn + UNPACK_CHUNK - 1) / UNPACK_CHUNK;
    if (bits_a == 0 || bits_b == 0) {
        J.zero = 1;
        *N_out = 0;
        return NBH_OK;
    }
    int bch = 0;
    int64_t N = 0;
    RCCHK(ntt_params(bits_a, bits_b, &bch, &N));
    if (bch < UNPACK_MIN_B) return fail(NBH_ERR_LENGTH, "product too large for device memory (chunk width below 32)");
    J.bch = bch;
    J.bmagic = ~0ull / (uint64_t)bch + 1;  // ceil(2^64 / bch)
    J.nca = (bits_a + bch - 1) / bch;
    J.ncb = (bits_b + bch - 1) / bch;
    J.nconv = J.nca + J.ncb - 1;
    *N_out = N;
    return NBH_OK;
}

// Sizing of a product against a cached spectrum: the spectrum's chunk width
// and length serve every a of up to bits_a_max bits.
int size_job_spec(DevJob &J, int64_t bits_a, const nbh_spectrum_s &S, int64_t *N_out) {
    J.rn = J.m + J.n;
    J.nchunks = (J.rn + UNPACK_CHUNK - 1) / UNPACK_CHUNK;
    if (bits_a > S.bits_a_max) return fail(NBH_ERR_ARG, "operand longer than the spectrum's bits_a_max");
    if (bits_a == 0 || S.zero) {
        J.zero = 1;
        *N_out = 0;
        return NBH_OK;
    }
    J.bch = S.bch;
    J.bmagic = ~0ull / (uint64_t)S.bch + 1;
    J.nca = (bits_a + S.bch - 1) / S.bch;
    J.ncb = S.ncb;
    J.nconv = J.nca + J.ncb - 1;
    J.bspec = 1;
    *N_out = S.N;
    return NBH_OK;
}

// A low short product keeps r[0..keep) only: the unpack stops at limb keep
// and its carry out of the top is dropped (keep = 0: the full m + n limbs).
void keep_limbs(DevJob &J, int64_t keep) {
    if (keep <= 0) return;
    J.rn = keep;
    J.nchunks = (keep + UNPACK_CHUNK - 1) / UNPACK_CHUNK;
}

// Twiddle doubles of a length-N plan (build_twiddles' table): (m-1) Q per
// stage of radix 3, 5, 7 or 4, none for radix 2.  Depends on N only.
int64_t twiddle_count(int64_t N) {
    int64_t span = N, m = N >> __builtin_ctzll((uint64_t)N), total = 0;
    for (int mf : {7, 5, 3, 3}) {
        if (m % mf) continue;
        total += (int64_t)(mf - 1) * (span / mf);
        span /= mf;
        m /= mf;
    }
    for (int64_t L = span; L >= 4; L >>= 2) total += 3 * (L / 4);
    return total;
}

// Geometry of a sized product in the one-workgroup kernel (k_small): half the
// block per prime, each half one thread per 4 points of a prime's two arrays
// (at least a wave, at most 512 threads); every stage its own pass when the
// half has a thread per 4 points of one array, else radix-4 pairs (then one
// thread per 16 points of both arrays suffices).  LDS for a's and b's arrays
// of both primes and both twiddle tables.  False when it does not fit one
// block: more threads than the kernel's bound, result segments than threads
// (the limb sums run in one round), staged Garner words than an array holds,
// loads than one round of four per thread, or LDS than a block has.
struct SmallGeo {
    int threads, ld, twl, nseg, cnt, single;
    size_t lds;
};
bool small_geo(const DevJob &J, int64_t N, SmallGeo *out) {
    if (J.zero || J.bspec || N < 16 || N > SMALL_TWC_MAX_N) return false;
    const int64_t half = std::min<int64_t>(SMALL_MAX_THREADS / 2, std::max<int64_t>(64, (2 * (N / 4) + 63) / 64 * 64));
    const int single = 4 * half >= N ? 1 : 0;
    const int64_t nt = 2 * half;
    const int64_t nseg = (J.rn + UNPACK_LPT - 1) / UNPACK_LPT;
    // the segment sums read the coefficients that start below limb 4 nseg
    const int64_t cnt = (64 * UNPACK_LPT * nseg + J.bch - 1) / J.bch + 1;
    const int64_t ld = lds_doubles((int)N);
    const int64_t twl = twiddle_count(N);
    const size_t lds = sizeof(double) * (size_t)(4 * ld + 2 * twl);
    const size_t lds_static = 2 * SMALL_MAX_STAGES * sizeof(StageDesc);
    int nst = 0;
    for (int64_t L = N; L > 1; nst++) L /= (L % 7 == 0 ? 7 : L % 5 == 0 ? 5 : L % 3 == 0 ? 3 : L % 4 == 0 ? 4 : 2);
    const int64_t sw = 2 * (int64_t)(sizeof(StageDesc) / sizeof(double)) * nst;
    if (nt > SMALL_MAX_THREADS || nseg > nt || cnt > ld || 2 * nseg > ld || lds + lds_static > g.lds_block ||
        nst > SMALL_MAX_STAGES || std::max(N, twl) > 4 * nt || sw > nt)
        return false;
    *out = SmallGeo{(int)nt, (int)ld, (int)twl, (int)nseg, (int)cnt, single, lds};
    return true;
}

// The products mul!/sqr! keep out of the NTT engine (Karatsuba and
// mul_basecase!, src/mul.jl:24-26,131,145,388): smaller operand below 64
// limbs or m + n below 304; squares below 160 limbs.
bool schoolbook_route(int64_t m, int64_t n, bool sqr) {
    return sqr ? m < SQR_FPNTT_THRESHOLD : std::min(m, n) < MUL_FPNTT_MIN || m + n < 2 * MUL_FPNTT_THRESHOLD;
}

int read_bits(const uint64_t *a, int64_t n, int64_t *bits, hipStream_t s = nullptr) {
    if (!s) s = g.stream;
    uint64_t top = 0;
    HIPCHK(hipMemcpyAsync(&top, a + (n - 1), sizeof top, hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    *bits = bits_from_top(n, top);
    return NBH_OK;
}

// One product on stream s (null: g.stream): pack+forward groups, mid,
// inverse groups, unpack.  keep > 0: only r[0..keep) is written, the product
// mod beta^keep.  The caller orders it with lib_begin / lib_end.
// sc: use that workspace and only cached plans; nothing is allocated (no
// device memory, no profiling events), and a length without a cached plan or
// a workspace need beyond sc->bytes is an error.
int run_single(uint64_t *r, const uint64_t *a, int64_t m, const uint64_t *b, int64_t n, int64_t bits_a,
               int64_t bits_b, bool sqr, const nbh_spectrum_s *spec = nullptr, int64_t keep = 0,
               hipStream_t s = nullptr, const Scratch *sc = nullptr) {
    if (!s) s = g.stream;
    if (spec) { b = spec->b; n = spec->n; bits_b = spec->bits_b; }
    if (!r || !a || m < 1 || (!sqr && (!b || n < 1))) return fail(NBH_ERR_ARG, "null buffer or empty operand");
    if (sqr) { b = a; n = m; }
    if (bits_a < 0) RCCHK(read_bits(a, m, &bits_a, s));
    if (sqr) bits_b = bits_a;
    else if (bits_b < 0) RCCHK(read_bits(b, n, &bits_b, s));
    DevJob J;
    std::memset(&J, 0, sizeof J);
    J.a = a;
    J.b = b;
    J.r = r;
    J.m = m;
    J.n = n;
    J.sqr = sqr ? 1 : 0;
    int64_t N = 0;
    if (spec) RCCHK(size_job_spec(J, bits_a, *spec, &N));
    else RCCHK(size_job(J, bits_a, bits_b, &N));
    keep_limbs(J, keep);
    if (J.zero) {
        HIPCHK(launch_fill_u64(r, J.rn, 0, s));
        return NBH_OK;
    }
    PlanEntry *P[2];
    // operands transformed here: a, and b unless it is a's square or cached
    const int nops = sqr || spec ? 1 : 2;
    const size_t xbytes = align256(sizeof(double) * (size_t)N);
    const size_t need = align256(sizeof(uint32_t) * (size_t)J.nchunks) + xbytes * 2 * nops;
    const bool prof = sc == nullptr;
    SmallGeo sg;
    if (small_geo(J, N, &sg)) {
        // the whole product in one workgroup: no workspace, one launch
        if (sc) {
            P[0] = find_plan(N, 0, true);
            P[1] = find_plan(N, 1, true);
            if (!P[0] || !P[1])
                return fail(NBH_ERR_ARG, "no cached plan for transform length " + std::to_string(N) +
                                             ": run one allocating product of this length (nbh_mul_dev) first");
        } else {
            RCCHK(get_plan(N, 0, &P[0], true));
            RCCHK(get_plan(N, 1, &P[1], true));
        }
        J.plan[0] = P[0]->dev;
        J.plan[1] = P[1]->dev;
        ProfScope ps(7, s, prof);
        HIPCHK(launch_small(J, sg.threads, sg.ld, sg.twl, sg.nseg, sg.cnt, sg.single, sg.lds, s));
        return NBH_OK;
    }
    char *w = nullptr;
    if (sc) {
        if (need > sc->bytes)
            return fail(NBH_ERR_ARG, "product needs " + std::to_string(need) + " bytes of workspace (transform length " +
                                         std::to_string(N) + "); the no-allocation entry point has " +
                                         std::to_string(sc->bytes));
        P[0] = find_plan(N, 0, true);
        P[1] = find_plan(N, 1, true);
        if (!P[0] || !P[1])
            return fail(NBH_ERR_ARG, "no cached plan for transform length " + std::to_string(N) +
                                         ": run one allocating product of this length (nbh_mul_dev) first");
        w = (char *)sc->p;
    } else {
        RCCHK(get_plan(N, 0, &P[0], true));
        RCCHK(get_plan(N, 1, &P[1], true));
        RCCHK(ensure_buf(&g.ws, &g.ws_size, need));
        w = (char *)g.ws;
    }
    J.chunk_flags = (uint32_t *)w;
    w += align256(sizeof(uint32_t) * (size_t)J.nchunks);
    for (int k = 0; k < 2; k++) {
        J.plan[k] = P[k]->dev;
        J.xa[k] = (double *)w;
        w += xbytes;
        if (spec) {
            J.xb[k] = spec->x[k];
        } else if (!sqr) {
            J.xb[k] = (double *)w;
            w += xbytes;
        }
    }
    // every launch carries the job record in its TileGen (no job array)
    const DevPlan &H = P[0]->host;
    const int K = H.ngroups;
    // few enough chunks: the unpack resolves the carries itself (no k_carry_fix)
    const bool lookback = J.nchunks > 1 && J.nchunks <= UNPACK_LB_MAX_CHUNKS;
    for (int gi = 0; gi + 1 < K; gi++) {
        TileGen tg{gi, nops, H.g[gi].tiles, gi == 0 ? TILE_PACK : 0, 0, J};
        ProfScope ps(0, s, prof);
        HIPCHK(launch_group(nullptr, nullptr, tg, 2 * nops * H.g[gi].tiles, 1, P[0]->grp_lds_bytes[gi],
                            P[0]->grp_threads[gi], s));
    }
    {
        TileGen tg{K - 1, 1, H.g[K - 1].tiles, (K == 1 ? TILE_PACK : 0) | (lookback ? TILE_CLEAR : 0), 0, J};
        ProfScope ps(1, s, prof);
        // a cached b still takes an LDS array, but only a is transformed
        HIPCHK(launch_mid(nullptr, nullptr, tg, 2 * H.g[K - 1].tiles, P[0]->mid_lds_bytes[sqr ? 1 : 0],
                          P[0]->mid_threads[sqr || spec ? 1 : 0], s));
    }
    for (int gi = K - 2; gi >= 0; gi--) {
        TileGen tg{gi, 1, H.g[gi].tiles, 0, 0, J};
        ProfScope ps(2, s, prof);
        HIPCHK(launch_group(nullptr, nullptr, tg, 2 * H.g[gi].tiles, 0, P[0]->grp_lds_bytes[gi],
                            P[0]->grp_threads[gi], s));
    }
    TileGen ug{0, 1, J.nchunks, 0, 0, J};
    {
        ProfScope ps(3, s, prof);
        HIPCHK(launch_unpack(nullptr, nullptr, ug, J.nchunks, J.bch, s, lookback));
    }
    if (J.nchunks > 1 && !lookback) {  // one chunk has no carry to pass on
        ProfScope ps(4, s, prof);
        HIPCHK(launch_carry_fix(nullptr, nullptr, ug, J.nchunks, s));
    }
    return NBH_OK;
}

// Low short product on stream s (mullo!/sqrlo!, src/mul.jl:305-378):
// r[0..k) = (a mod beta^k) * (b mod beta^k) mod beta^k.  The clamped product
// takes the engine mul!/sqr! would give it (every engine is exact, so the low
// limbs are the reference's whichever it takes), with only its low k limbs
// unpacked or summed.
int run_lo(uint64_t *r, const uint64_t *a, int64_t la, const uint64_t *b, int64_t lb, int64_t k, int64_t bits_a,
           int64_t bits_b, bool sqr, hipStream_t s = nullptr) {
    if (!s) s = g.stream;
    if (!r || !a || la < 1 || k < 1 || (!sqr && (!b || lb < 1)))
        return fail(NBH_ERR_ARG, "null buffer, empty operand or k < 1");
    if (sqr) { b = a; lb = la; bits_b = bits_a; }
    la = std::min(la, k);
    lb = std::min(lb, k);
    // a magnitude bound of the whole operand bounds its low limbs too
    if (bits_a >= 0) bits_a = std::min(bits_a, 64 * la);
    if (bits_b >= 0) bits_b = std::min(bits_b, 64 * lb);
    if (schoolbook_route(la, lb, sqr)) {
        const bool swap = lb > la;
        const BaseJob j{swap ? b : a, swap ? a : b, r, std::max(la, lb), std::min(la, lb), k};
        HIPCHK(launch_basecase_one(j, s));
        return NBH_OK;
    }
    return run_single(r, a, la, sqr ? nullptr : b, lb, bits_a, bits_b, sqr, nullptr, k, s);
}

// The single-product resources (workspace, I/O and pinned buffers) serve one call at a time, in submission order, whatever stream a
// call runs on: each call's stream first waits for the event recorded at the
// end of the previous call (lib_begin) and records it again at its own end
// (lib_end).  Single products on a caller's stream run on that stream, so
// the caller's own work before and after them needs no cross-stream wait.
int lib_begin(hipStream_t s) {
    // the library's own stream is already ordered after its own last call
    // (a caller's stream could be a new one at a destroyed one's address, so
    // it always waits)
    if (g.last_valid && !(s == g.stream && g.last_on_lib)) HIPCHK(hipStreamWaitEvent(s, g.last, 0));
    return NBH_OK;
}
int lib_end(hipStream_t s) {
    HIPCHK(hipEventRecord(g.last, s));
    g.last_valid = true;
    g.last_on_lib = s == g.stream;
    return NBH_OK;
}

// For entry points that run on g.stream: order it after `user`'s prior work
// and the previous call; after the call, make `user` wait for ours.
int stream_enter(void *user) {
    RCCHK(lib_begin(g.stream));
    if (!user) return NBH_OK;
    HIPCHK(hipEventRecord(g.ev_in, (hipStream_t)user));
    HIPCHK(hipStreamWaitEvent(g.stream, g.ev_in, 0));
    return NBH_OK;
}
int stream_leave(void *user) {
    RCCHK(lib_end(g.stream));
    if (!user) return NBH_OK;
    HIPCHK(hipEventRecord(g.ev_out, g.stream));
    HIPCHK(hipStreamWaitEvent((hipStream_t)user, g.ev_out, 0));
    return NBH_OK;
}

// A single product (or short product) on the caller's stream.
int run_on(void *stream, const std::function<int(hipStream_t)> &f) {
    const hipStream_t s = stream ? (hipStream_t)stream : g.stream;
    RCCHK(lib_begin(s));
    const int rc = f(s);
    const int rc2 = lib_end(s);
    return rc != NBH_OK ? rc : rc2;
}

// Whether a host-pointer product (m, n already clamped to keep) runs whole in
// the one-workgroup kernel.
bool host_small(int64_t m, int64_t n, int64_t bits_a, int64_t bits_b, bool sqr, int64_t keep) {
    if (keep && schoolbook_route(m, n, sqr)) return false;
    DevJob J;
    std::memset(&J, 0, sizeof J);
    J.m = m;
    J.n = n;
    J.sqr = sqr ? 1 : 0;
    int64_t N = 0;
    if (size_job(J, bits_a, bits_b, &N) != NBH_OK) return false;
    keep_limbs(J, keep);
    SmallGeo sg;
    return small_geo(J, N, &sg);
}

// Host buffers in and out; keep > 0 is the low short product r[0..keep).
int host_product(uint64_t *r, const uint64_t *a, int64_t m, const uint64_t *b, int64_t n, bool sqr,
                 int64_t keep = 0) {
    if (!r || !a || m < 1 || (!sqr && (!b || n < 1)) || keep < 0)
        return fail(NBH_ERR_ARG, "null buffer, empty operand or negative k");
    if (sqr) n = m;
    if (keep) {
        m = std::min(m, keep);
        n = std::min(n, keep);
    }
    RCCHK(lib_begin(g.stream));
    const int64_t bits_a = bits_from_top(m, a[m - 1]);
    const int64_t bits_b = sqr ? bits_a : bits_from_top(n, b[n - 1]);
    const int64_t rl = keep ? keep : m + n;
    const size_t sa = align256(8 * (size_t)m), sb = sqr ? 0 : align256(8 * (size_t)n);
    const size_t sr = align256(8 * (size_t)rl);
    RCCHK(ensure_buf(&g.io, &g.io_size, sa + sb + sr));
    uint64_t *da = (uint64_t *)g.io;
    uint64_t *db = (uint64_t *)((char *)g.io + sa);
    uint64_t *dr = (uint64_t *)((char *)g.io + sa + sb);
    const bool pinned = sa + sb + sr <= PIN_MAX_BYTES;
    // the one-workgroup kernel reads each limb once, in one round of loads:
    // over the bus that costs less than two copy commands at any size it takes
    if (pinned && (sa + sb + sr >= ZERO_COPY_MIN_BYTES || host_small(m, n, bits_a, bits_b, sqr, keep))) {
        RCCHK(ensure_pinned(sa + sb + sr));
        if (g.pin_dev) {
            // the kernels read the operands from and write r to the pinned
            // buffer over the bus: no copy commands
            char *p = (char *)g.pin;
            std::memcpy(p, a, 8 * (size_t)m);
            if (!sqr) std::memcpy(p + sa, b, 8 * (size_t)n);
            char *d = (char *)g.pin_dev;
            uint64_t *za = (uint64_t *)d, *zb = (uint64_t *)(d + sa), *zr = (uint64_t *)(d + sa + sb);
            if (keep) RCCHK(run_lo(zr, za, m, sqr ? nullptr : zb, n, keep, bits_a, bits_b, sqr));
            else RCCHK(run_single(zr, za, m, sqr ? nullptr : zb, n, bits_a, bits_b, sqr));
            RCCHK(lib_end(g.stream));
            HIPCHK(hipStreamSynchronize(g.stream));
            std::memcpy(r, p + sa + sb, 8 * (size_t)rl);
            return NBH_OK;
        }
    }
    if (pinned) {
        // operands into the pinned buffer with the device layout, one DMA in, one out
        RCCHK(ensure_pinned(sa + sb + sr));
        char *p = (char *)g.pin;
        std::memcpy(p, a, 8 * (size_t)m);
        if (!sqr) std::memcpy(p + sa, b, 8 * (size_t)n);
        HIPCHK(hipMemcpyAsync(da, p, sa + sb, hipMemcpyHostToDevice, g.stream));
    } else {
        HIPCHK(hipMemcpyAsync(da, a, 8 * (size_t)m, hipMemcpyHostToDevice, g.stream));
        if (!sqr) HIPCHK(hipMemcpyAsync(db, b, 8 * (size_t)n, hipMemcpyHostToDevice, g.stream));
    }
    if (keep) RCCHK(run_lo(dr, da, m, sqr ? nullptr : db, n, keep, bits_a, bits_b, sqr));
    else RCCHK(run_single(dr, da, m, sqr ? nullptr : db, n, bits_a, bits_b, sqr));
    if (pinned) {
        char *pr = (char *)g.pin + sa + sb;
        HIPCHK(hipMemcpyAsync(pr, dr, 8 * (size_t)rl, hipMemcpyDeviceToHost, g.stream));
        RCCHK(lib_end(g.stream));
        HIPCHK(hipStreamSynchronize(g.stream));
        std::memcpy(r, pr, 8 * (size_t)rl);
    } else {
        HIPCHK(hipMemcpyAsync(r, dr, 8 * (size_t)rl, hipMemcpyDeviceToHost, g.stream));
        RCCHK(lib_end(g.stream));
        HIPCHK(hipStreamSynchronize(g.stream));
    }
    return NBH_OK;
}

}  // namespace

// ---------------------------------------------------------------------------
// Batches

struct nbh_batch_s {
    struct Phase {
        int kind;        // 0 group, 1 mid, 2 unpack, 3 carry fix, 5 schoolbook products
        int fwd;
        int64_t off, count;
        size_t lds;
        int threads;
        int step;        // launches of one step are independent of each other
    };
    std::vector<DevJob> jobs;
    std::vector<TileRange> ranges;  // host copies (kept alive for async uploads)
    std::vector<BaseJob> base;      // products below the NTT thresholds (schoolbook)
    BaseJob *dbase = nullptr;
    int64_t ntiles = 0;
    DevJob *djobs = nullptr;
    Tile *dtiles = nullptr;
    void *ws = nullptr;
    bool external = false;       // tables and workspace live in a caller's arena
    std::vector<Phase> phases;
    int nsteps = 0;
    std::vector<hipEvent_t> ev;  // fork event per step + one join event per launch
    int nside = -1;              // side streams for this batch's steps (-1: the library setting)
    hipGraph_t graph = nullptr;  // the launch sequence captured once (nbh_batch_capture)
    hipGraphExec_t gexec = nullptr;
};

namespace {

void batch_free(nbh_batch_s *B) {
    if (!B) return;
    if (B->gexec) (void)hipGraphExecDestroy(B->gexec);
    if (B->graph) (void)hipGraphDestroy(B->graph);
    if (!B->external) {
        if (B->djobs) (void)hipFree(B->djobs);
        if (B->dtiles) (void)hipFree(B->dtiles);
        if (B->ws) (void)hipFree(B->ws);
        if (B->dbase) (void)hipFree(B->dbase);
    }
    for (auto e : B->ev) g.ev_pool.push_back(e);
    delete B;
}

// Plan a batch.  With `need` set, only report the bytes the batch occupies in
// an arena (workspace + job and tile tables) and create nothing.  With `arena`
// set, place everything inside it and upload the tables on `astream` -- no
// device allocation and no host synchronisation, so a server can plan batches
// while others run.  Otherwise allocate with hipMalloc.
// A job as the planner sees it: a product, a square (b == NULL) or a product
// against a cached spectrum (spec set; b and n are the spectrum's).
struct JobIn {
    uint64_t *r;
    const uint64_t *a, *b;
    int64_t m, n, bits_a, bits_b;
    const nbh_spectrum_s *spec;
    int64_t keep;  // > 0: low short product r[0..keep) (a, and b unless cached, clamped to keep limbs)
};

int batch_plan(const JobIn *jobs, int64_t count, nbh_batch_s **out, size_t *need, void *arena,
               size_t arena_bytes, hipStream_t astream) {
    if (count < 1 || !jobs) return fail(NBH_ERR_ARG, "empty batch");
    std::unique_ptr<nbh_batch_s, void (*)(nbh_batch_s *)> B(new nbh_batch_s, batch_free);
    std::vector<DevJob> &J = B->jobs;
    J.resize((size_t)count);
    std::vector<int64_t> bits((size_t)(2 * count), 0);
    std::vector<char> isbase((size_t)count, 0);
    bool need_top = false;
    for (int64_t i = 0; i < count; i++) {
        const JobIn &q = jobs[i];
        DevJob &d = J[(size_t)i];
        std::memset(&d, 0, sizeof d);
        const bool sqr = q.b == nullptr && !q.spec;
        if (!q.r || !q.a || q.m < 1 || (!sqr && q.n < 1) || q.keep < 0)
            return fail(NBH_ERR_ARG, "bad job " + std::to_string(i));
        d.a = q.a;
        d.b = sqr ? q.a : q.b;
        d.r = q.r;
        d.m = q.m;
        d.n = sqr ? q.m : q.n;
        d.sqr = sqr;
        int64_t ba = q.bits_a, bb = sqr ? q.bits_a : q.bits_b;
        if (q.keep > 0) {
            // limbs at or above keep cannot reach r (mullo!, src/mul.jl:308-309);
            // a cached spectrum's operand stays whole
            d.m = std::min(d.m, q.keep);
            if (!q.spec) d.n = std::min(d.n, q.keep);
            if (ba > 0) ba = std::min(ba, 64 * d.m);
            if (bb > 0 && !q.spec) bb = std::min(bb, 64 * d.n);
        }
        bits[(size_t)(2 * i)] = ba;
        bits[(size_t)(2 * i + 1)] = bb;
        if (ba <= 0 || (!sqr && !q.spec && bb <= 0)) need_top = true;
        // below the reference's NTT thresholds: schoolbook
        if (!q.spec && schoolbook_route(d.m, d.n, sqr)) {
            isbase[(size_t)i] = 1;
            const bool swap = d.n > d.m;
            B->base.push_back(BaseJob{swap ? d.b : d.a, swap ? d.a : d.b, d.r, std::max(d.m, d.n),
                                      std::min(d.m, d.n), q.keep > 0 ? q.keep : d.m + d.n});
        }
    }
    // schoolbook products need no sizing, plans or workspace
    if (need_top) {
        need_top = false;
        for (int64_t i = 0; i < count; i++)
            if (!isbase[(size_t)i] && (bits[(size_t)(2 * i)] <= 0 ||
                                       (jobs[i].b && !jobs[i].spec && bits[(size_t)(2 * i + 1)] <= 0)))
                need_top = true;
    }
    if (need_top && (need || arena)) return fail(NBH_ERR_ARG, "arena batches need the operands' bit lengths");
    if (need_top) {
        DevJob *tmp = nullptr;
        uint64_t *tops = nullptr;
        HIPCHK(hipMalloc((void **)&tmp, sizeof(DevJob) * (size_t)count));
        HIPCHK(hipMalloc((void **)&tops, 16 * (size_t)count));
        HIPCHK(hipMemcpyAsync(tmp, J.data(), sizeof(DevJob) * (size_t)count, hipMemcpyHostToDevice, g.stream));
        HIPCHK(launch_top_limbs(tmp, count, tops, g.stream));
        std::vector<uint64_t> h((size_t)(2 * count));
        HIPCHK(hipMemcpyAsync(h.data(), tops, 16 * (size_t)count, hipMemcpyDeviceToHost, g.stream));
        HIPCHK(hipStreamSynchronize(g.stream));
        (void)hipFree(tmp);
        (void)hipFree(tops);
        for (int64_t i = 0; i < count; i++) {
            if (bits[(size_t)(2 * i)] <= 0) bits[(size_t)(2 * i)] = bits_from_top(J[(size_t)i].m, h[(size_t)(2 * i)]);
            if (bits[(size_t)(2 * i + 1)] <= 0)
                bits[(size_t)(2 * i + 1)] = bits_from_top(J[(size_t)i].n, h[(size_t)(2 * i + 1)]);
        }
    }
    // size every job, get plans, lay out the workspace
    std::vector<PlanEntry *> plan0((size_t)count, nullptr), plan1((size_t)count, nullptr);
    std::vector<int64_t> Ns((size_t)count, 0);
    size_t ws = 256;  // offsets start past a reserved block: offset 0 would read as "no buffer"
    std::vector<size_t> offs((size_t)count);
    for (int64_t i = 0; i < count; i++) {
        DevJob &d = J[(size_t)i];
        offs[(size_t)i] = ws;
        if (isbase[(size_t)i]) {
            d.zero = 1;  // no NTT work; the schoolbook phase writes r
            continue;
        }
        if (jobs[i].spec) RCCHK(size_job_spec(d, bits[(size_t)(2 * i)], *jobs[i].spec, &Ns[(size_t)i]));
        else RCCHK(size_job(d, bits[(size_t)(2 * i)], bits[(size_t)(2 * i + 1)], &Ns[(size_t)i]));
        keep_limbs(d, jobs[i].keep);
        offs[(size_t)i] = ws;
        ws += align256(sizeof(uint32_t) * (size_t)d.nchunks);
        if (d.zero) continue;
        RCCHK(get_plan(Ns[(size_t)i], 0, &plan0[(size_t)i]));
        RCCHK(get_plan(Ns[(size_t)i], 1, &plan1[(size_t)i]));
        ws += align256(sizeof(double) * (size_t)Ns[(size_t)i]) * (d.sqr || d.bspec ? 2 : 4);
    }
    B->ws = nullptr;
    std::vector<size_t> spec_off((size_t)count);
    for (int64_t i = 0; i < count; i++) spec_off[(size_t)i] = offs[(size_t)i];
    for (int64_t i = 0; i < count; i++) {
        DevJob &d = J[(size_t)i];
        char *w = reinterpret_cast<char *>((uintptr_t)spec_off[(size_t)i]);  // offset, rebased below
        d.chunk_flags = (uint32_t *)w;
        w += align256(sizeof(uint32_t) * (size_t)d.nchunks);
        if (d.zero) continue;
        const size_t xb = align256(sizeof(double) * (size_t)Ns[(size_t)i]);
        d.plan[0] = plan0[(size_t)i]->dev;
        d.plan[1] = plan1[(size_t)i]->dev;
        for (int k = 0; k < 2; k++) {
            d.xa[k] = (double *)w;
            w += xb;
            if (d.bspec) {
                d.xb[k] = jobs[i].spec->x[k];  // a real address: not rebased
            } else if (!d.sqr) {
                d.xb[k] = (double *)w;
                w += xb;
            }
        }
    }
    // phase tile lists; within a phase, bigger tiles first
    std::vector<TileRange> &ranges = B->ranges;
    int64_t &ntiles = B->ntiles;
    int maxK = 0;
    for (int64_t i = 0; i < count; i++)
        if (!J[(size_t)i].zero) maxK = std::max(maxK, plan0[(size_t)i]->host.ngroups);
    // One launch per (phase, thread class): tiles of very different sizes get
    // their own block shape.  Within a launch, bigger tiles first, then all of
    // p1's tiles before p2's: the blocks resident on a CU at a time then mostly
    // share one prime's twiddle tables in its caches (batch 1% faster).
    struct TileReq {
        size_t lds;
        int threads;
        Tile t;          // first tile of the range
        uint32_t count;
    };
    auto push_phase = [&](int kind, int fwd, std::vector<TileReq> &list) {
        if (list.empty()) return;
        std::stable_sort(list.begin(), list.end(), [](const TileReq &x, const TileReq &y) {
            return x.threads != y.threads ? x.threads > y.threads : x.lds != y.lds ? x.lds > y.lds : x.t.prime < y.t.prime;
        });
        // LDS tiles of one block size share a launch when the largest LDS
        // request keeps as many blocks per CU as the smaller one would (and,
        // for strided groups, the same build): the kernels take their tile
        // layout from the plan, not from the launch's LDS size, and the
        // smaller tiles then fill the tail of the big launch instead of
        // queueing launches of their own.  Column groups (marker LDS sizes)
        // and the unpack stay one launch per class.
        auto joins = [&](const nbh_batch_s::Phase &ph, size_t lds, int threads) {
            // mid tiles also join a launch of larger blocks: their extra threads
            // idle, and they take the general load/store passes (the edge
            // passes need exactly one thread per 16 points)
            if (kind == 1)
                return lds <= ph.lds && threads <= ph.threads &&
                       lds_blocks_per_cu(false, ph.lds, ph.threads) >= lds_blocks_per_cu(false, lds, threads);
            if (threads != ph.threads) return false;
            if (lds == ph.lds) return true;
            if (kind != 0 || ph.lds <= COL_BIG || lds <= COL_BIG) return false;
            return group_six_waves(ph.lds, ph.threads) == group_six_waves(lds, ph.threads) &&
                   lds_blocks_per_cu(true, ph.lds, ph.threads) >= lds_blocks_per_cu(true, lds, ph.threads);
        };
        size_t i = 0;
        const int step = B->nsteps++;
        while (i < list.size()) {
            nbh_batch_s::Phase ph{kind, fwd, ntiles, 0, list[i].lds, list[i].threads, step};
            while (i < list.size() && joins(ph, list[i].lds, list[i].threads)) {
                ranges.push_back(TileRange{list[i].t, list[i].count, (uint32_t)ntiles});
                ntiles += list[i].count;
                ph.count += list[i].count;
                i++;
            }
            if (kind >= 3) ph.lds = 0;
            B->phases.push_back(ph);
        }
        list.clear();
    };
    std::vector<TileReq> list;
    auto add_group_tiles = [&](int64_t i, int gi, int nops, uint8_t flags, size_t lds, int threads) {
        const DevPlan &H = plan0[(size_t)i]->host;
        for (int k = 0; k < 2; k++)
            for (int op = 0; op < nops; op++)
                list.push_back({lds, threads, Tile{(uint32_t)i, (uint8_t)k, (uint8_t)op, (uint8_t)gi, flags, 0, 0},
                                (uint32_t)H.g[gi].tiles});
    };
    for (int d = maxK - 1; d >= 1; d--) {  // forward strided groups, farthest from the mid first
        for (int64_t i = 0; i < count; i++) {
            if (J[(size_t)i].zero) continue;
            const int K = plan0[(size_t)i]->host.ngroups;
            const int gi = K - 1 - d;
            if (gi < 0) continue;
            add_group_tiles(i, gi, J[(size_t)i].sqr || J[(size_t)i].bspec ? 1 : 2, gi == 0 ? TILE_PACK : 0,
                            plan0[(size_t)i]->grp_lds_bytes[gi], plan0[(size_t)i]->grp_threads[gi]);
        }
        push_phase(0, 1, list);
    }
    for (int64_t i = 0; i < count; i++) {
        if (J[(size_t)i].zero) continue;
        const int K = plan0[(size_t)i]->host.ngroups;
        add_group_tiles(i, K - 1, 1, K == 1 ? TILE_PACK : 0, plan0[(size_t)i]->mid_lds_bytes[J[(size_t)i].sqr ? 1 : 0],
                        plan0[(size_t)i]->mid_threads[J[(size_t)i].sqr || J[(size_t)i].bspec ? 1 : 0]);
    }
    push_phase(1, 1, list);
    for (int d = 1; d <= maxK - 1; d++) {
        for (int64_t i = 0; i < count; i++) {
            if (J[(size_t)i].zero) continue;
            const int K = plan0[(size_t)i]->host.ngroups;
            const int gi = K - 1 - d;
            if (gi < 0) continue;
            add_group_tiles(i, gi, 1, 0, plan0[(size_t)i]->grp_lds_bytes[gi], plan0[(size_t)i]->grp_threads[gi]);
        }
        push_phase(0, 0, list);
    }
    std::vector<TileReq> list2;
    // one unpack launch: its field `lds` carries the smallest chunk width (zero
    // products and schoolbook jobs have none)
    int bmin = 64;
    for (int64_t i = 0; i < count; i++)
        if (!J[(size_t)i].zero) bmin = std::min(bmin, (int)J[(size_t)i].bch);
    for (int64_t i = 0; i < count; i++) {
        const uint32_t nch = (uint32_t)J[(size_t)i].nchunks;
        list.push_back({(size_t)bmin, UNPACK_THREADS, Tile{(uint32_t)i, 0, 0, 0, 0, 0, 0}, nch});
        // carries pass into chunks 1.. only; a batch of one-chunk products has no carry launch
        if (nch > 1) list2.push_back({0, UNPACK_THREADS, Tile{(uint32_t)i, 0, 0, 0, 0, 1, 0}, nch - 1});
    }
    push_phase(2, 0, list);
    push_phase(3, 0, list2);
    // the schoolbook products: one launch, beside the first NTT step; one
    // workgroup per product, the most limb products first so the longest
    // blocks are dispatched in the first round (products are independent)
    if (!B->base.empty()) {
        std::stable_sort(B->base.begin(), B->base.end(), [](const BaseJob &x, const BaseJob &y) {
            return x.m * std::min(x.n, x.rn) > y.m * std::min(y.n, y.rn);
        });
        const int step = B->phases.empty() ? B->nsteps++ : B->phases.front().step;
        B->phases.insert(B->phases.begin(), nbh_batch_s::Phase{5, 0, 0, (int64_t)B->base.size(), 0,
                                                              UNPACK_THREADS, step});
    }
    // workspace offsets were laid out from 0: rebase them onto the real workspace
    const size_t ws_bytes = align256(std::max<size_t>(ws, 256));
    const size_t jobs_bytes = align256(sizeof(DevJob) * (size_t)count);
    // tile list followed by the ranges it is expanded from
    const size_t tiles_list_bytes = align256(sizeof(Tile) * (size_t)std::max<int64_t>(ntiles, 1));
    const size_t tiles_bytes = tiles_list_bytes + align256(sizeof(TileRange) * std::max<size_t>(ranges.size(), 1));
    const size_t base_bytes = B->base.empty() ? 0 : align256(sizeof(BaseJob) * B->base.size());
    if (need) {
        *need = ws_bytes + jobs_bytes + tiles_bytes + base_bytes;
        return NBH_OK;
    }
    if (arena) {
        if (ws_bytes + jobs_bytes + tiles_bytes + base_bytes > arena_bytes)
            return fail(NBH_ERR_BUDGET, "arena too small");
        B->external = true;
        B->ws = arena;
        B->djobs = (DevJob *)((char *)arena + ws_bytes);
        B->dtiles = (Tile *)((char *)arena + ws_bytes + jobs_bytes);
        if (base_bytes) B->dbase = (BaseJob *)((char *)arena + ws_bytes + jobs_bytes + tiles_bytes);
    } else {
        HIPCHK(hipMalloc(&B->ws, ws_bytes));
        HIPCHK(hipMalloc((void **)&B->djobs, jobs_bytes));
        HIPCHK(hipMalloc((void **)&B->dtiles, tiles_bytes));
        if (base_bytes) HIPCHK(hipMalloc((void **)&B->dbase, base_bytes));
    }
    TileRange *dranges = (TileRange *)((char *)B->dtiles + tiles_list_bytes);
    auto rebase = [&](auto *p) {
        return p ? (decltype(p))((char *)B->ws + reinterpret_cast<uintptr_t>(p)) : p;
    };
    for (auto &d : J) {
        d.chunk_flags = rebase(d.chunk_flags);
        for (int k = 0; k < 2; k++) {
            d.xa[k] = rebase(d.xa[k]);
            if (!d.bspec) d.xb[k] = rebase(d.xb[k]);
        }
    }
    if (arena) {
        HIPCHK(hipMemcpyAsync(B->djobs, J.data(), sizeof(DevJob) * (size_t)count, hipMemcpyHostToDevice, astream));
        HIPCHK(hipMemcpyAsync(dranges, ranges.data(), sizeof(TileRange) * ranges.size(), hipMemcpyHostToDevice,
                              astream));
        HIPCHK(launch_expand_tiles(dranges, (int64_t)ranges.size(), B->dtiles, astream));
        if (base_bytes)
            HIPCHK(hipMemcpyAsync(B->dbase, B->base.data(), sizeof(BaseJob) * B->base.size(), hipMemcpyHostToDevice,
                                  astream));
    } else {
        HIPCHK(hipMemcpy(B->djobs, J.data(), sizeof(DevJob) * (size_t)count, hipMemcpyHostToDevice));
        HIPCHK(hipMemcpy(dranges, ranges.data(), sizeof(TileRange) * ranges.size(), hipMemcpyHostToDevice));
        HIPCHK(launch_expand_tiles(dranges, (int64_t)ranges.size(), B->dtiles, g.stream));
        if (base_bytes)
            HIPCHK(hipMemcpy(B->dbase, B->base.data(), sizeof(BaseJob) * B->base.size(), hipMemcpyHostToDevice));
        HIPCHK(hipStreamSynchronize(g.stream));
    }
    B->ev.resize((size_t)B->nsteps + B->phases.size(), nullptr);
    for (auto &e : B->ev) {
        if (!g.ev_pool.empty()) {
            e = g.ev_pool.back();
            g.ev_pool.pop_back();
        } else {
            HIPCHK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
        }
    }
    *out = B.release();
    return NBH_OK;
}

int batch_create(const nbh_job *jobs, int64_t count, nbh_batch_s **out, size_t *need = nullptr,
                 void *arena = nullptr, size_t arena_bytes = 0, hipStream_t astream = nullptr) {
    if (count < 1 || !jobs) return fail(NBH_ERR_ARG, "empty batch");
    std::vector<JobIn> in((size_t)count);
    for (int64_t i = 0; i < count; i++) {
        const nbh_job &q = jobs[i];
        in[(size_t)i] = JobIn{q.r, q.a, q.b, q.m, q.n, q.bits_a, q.bits_b, nullptr, q.k};
    }
    return batch_plan(in.data(), count, out, need, arena, arena_bytes, astream);
}

int launch_phase(nbh_batch_s *B, const nbh_batch_s::Phase &ph, hipStream_t s) {
    const TileGen none{0, 1, 1, 0, 0};
    const Tile *t = B->dtiles + ph.off;
    ProfScope ps(ph.kind == 0 ? (ph.fwd ? 0 : 2) : ph.kind == 1 ? 1 : ph.kind == 2 ? 3 : ph.kind == 5 ? 6 : 4, s);
    switch (ph.kind) {
    case 5: HIPCHK(launch_basecase(B->dbase, ph.count, s)); break;
    case 0: HIPCHK(launch_group(B->djobs, t, none, ph.count, ph.fwd, ph.lds, ph.threads, s)); break;
    case 1: HIPCHK(launch_mid(B->djobs, t, none, ph.count, ph.lds, ph.threads, s)); break;
    case 2: HIPCHK(launch_unpack(B->djobs, t, none, ph.count, (int)ph.lds, s)); break;  // lds: smallest b
    default: HIPCHK(launch_carry_fix(B->djobs, t, none, ph.count, s)); break;
    }
    return NBH_OK;
}

// A step's launches (one per block shape) are independent: they fork onto the
// library's side streams after an event on `s` and join back before the next
// step, so small shape classes run beside the big ones instead of after them.
int batch_run(nbh_batch_s *B, hipStream_t s) {
    const int nside = B->nside >= 0 ? B->nside : g.nside;
    size_t i = 0;
    while (i < B->phases.size()) {
        size_t j = i;
        while (j < B->phases.size() && B->phases[j].step == B->phases[i].step) j++;
        if (j - i == 1 || nside == 0) {
            for (size_t k = i; k < j; k++) RCCHK(launch_phase(B, B->phases[k], s));
        } else {
            hipEvent_t fork = B->ev[(size_t)B->phases[i].step];
            HIPCHK(hipEventRecord(fork, s));
            for (size_t k = i; k < j; k++) {
                hipStream_t side = g.side[(k - i) % nside];
                HIPCHK(hipStreamWaitEvent(side, fork, 0));
                RCCHK(launch_phase(B, B->phases[k], side));
                hipEvent_t done = B->ev[(size_t)B->nsteps + k];
                HIPCHK(hipEventRecord(done, side));
                HIPCHK(hipStreamWaitEvent(s, done, 0));
            }
        }
        i = j;
    }
    return NBH_OK;
}

}  // namespace

// ---------------------------------------------------------------------------
// Batched 2x2 cofactor steps: four products per instance in one product
// batch, then one combine launch for the sums / signed differences.

struct nbh_hgcd_s {
    nbh_batch_s *products = nullptr;
    void *tmp = nullptr;
    CombineJob *dcomb = nullptr;
    int64_t ncomb = 0;
    bool lens = false;  // combine jobs carry normlen outputs (a level)
};

namespace {

// One combine workgroup walks its whole output, so the longest outputs set the
// launch's tail: they go first, to be dispatched in the first round of blocks.
// Jobs are independent (own outputs, sign and length words), so order is free.
void sort_combine_jobs(std::vector<CombineJob> &comb) {
    std::stable_sort(comb.begin(), comb.end(),
                     [](const CombineJob &x, const CombineJob &y) { return x.lr > y.lr; });
}

void hgcd_free(nbh_hgcd_s *H) {
    if (!H) return;
    batch_free(H->products);
    if (H->tmp) (void)hipFree(H->tmp);
    if (H->dcomb) (void)hipFree(H->dcomb);
    delete H;
}

int hgcd_create(const nbh_hgcd_job *jobs, int64_t count, int mode, nbh_hgcd_s **out) {
    if (!jobs || count < 1 || (mode != NBH_HGCD_FORWARD && mode != NBH_HGCD_ADJUST && mode != NBH_HGCD_ADJUST_FULL))
        return fail(NBH_ERR_ARG, "bad cofactor-step batch");
    const bool full = mode == NBH_HGCD_ADJUST_FULL;
    std::unique_ptr<nbh_hgcd_s, void (*)(nbh_hgcd_s *)> H(new nbh_hgcd_s, hgcd_free);
    size_t words = 0;
    for (int64_t i = 0; i < count; i++) {
        const nbh_hgcd_job &q = jobs[i];
        if (q.mn < 1 || q.n < 1 || !q.m00 || !q.m01 || !q.m10 || !q.m11 || !q.u || !q.v || !q.r0 || !q.r1 ||
            (mode != NBH_HGCD_FORWARD && (!q.s0 || !q.s1)) || (full && (!q.ra || !q.rb || q.nn < 1)))
            return fail(NBH_ERR_ARG, "bad cofactor-step instance " + std::to_string(i));
        words += 4 * (size_t)(q.n + q.mn);
    }
    HIPCHK(hipMalloc(&H->tmp, 8 * words));
    std::vector<nbh_job> prods;
    std::vector<CombineJob> comb;
    uint64_t *t = (uint64_t *)H->tmp;
    for (int64_t i = 0; i < count; i++) {
        const nbh_hgcd_job &q = jobs[i];
        const int64_t lp = q.n + q.mn;
        uint64_t *P[4] = {t, t + lp, t + 2 * lp, t + 3 * lp};
        t += 4 * lp;
        // forward: m00 u, m01 v, m10 u, m11 v; adjust: m11 u, m01 v, m00 v, m10 u
        const uint64_t *mat[4] = {q.m00, q.m01, q.m10, q.m11};
        const uint64_t *vec[4] = {q.u, q.v, q.u, q.v};
        if (mode != NBH_HGCD_FORWARD) {
            mat[0] = q.m11; vec[0] = q.u;
            mat[1] = q.m01; vec[1] = q.v;
            mat[2] = q.m00; vec[2] = q.v;
            mat[3] = q.m10; vec[3] = q.u;
        }
        for (int k = 0; k < 4; k++) prods.push_back(nbh_job{P[k], vec[k], mat[k], q.n, q.mn, -1, -1, 0});
        const int sub = mode != NBH_HGCD_FORWARD;
        if (full) {
            // hgcd_matrix_adjust! (src/gcd.jl:398-437): p = n, L = p + max(mn + 1, nn) + 1
            const int64_t L = q.n + std::max(q.mn + 1, q.nn) + 1;
            comb.push_back(CombineJob{P[0], P[1], lp, lp, q.r0, L, q.s0, 1, 0, q.ra, q.nn, q.n});
            comb.push_back(CombineJob{P[2], P[3], lp, lp, q.r1, L, q.s1, 1, 0, q.rb, q.nn, q.n});
        } else {
            const int64_t lr = sub ? lp : lp + 1;
            comb.push_back(CombineJob{P[0], P[1], lp, lp, q.r0, lr, q.s0, sub, 0, nullptr, 0, 0});
            comb.push_back(CombineJob{P[2], P[3], lp, lp, q.r1, lr, q.s1, sub, 0, nullptr, 0, 0});
        }
    }
    RCCHK(batch_create(prods.data(), (int64_t)prods.size(), &H->products));
    sort_combine_jobs(comb);
    H->ncomb = (int64_t)comb.size();
    HIPCHK(hipMalloc((void **)&H->dcomb, sizeof(CombineJob) * comb.size()));
    HIPCHK(hipMemcpy(H->dcomb, comb.data(), sizeof(CombineJob) * comb.size(), hipMemcpyHostToDevice));
    *out = H.release();
    return NBH_OK;
}

// One hgcd! level (src/gcd.jl:546-554) per instance: hgcd_matrix_adjust!
// (src/gcd.jl:398-437) on a, b with M1 and hgcd_matrix_mul! (src/gcd.jl:
// 372-391) M <- M * M1.  The adjust reads M1 and the low p limbs of a, b; the
// matrix product reads M and M1: disjoint from what the other writes, so all
// twelve products of an instance go into one product batch (into a private
// workspace) and the six outputs are written in place by one combine launch
// after every product is done.
int hgcd_level_create(const nbh_hgcd_level_job *jobs, int64_t count, nbh_hgcd_s **out) {
    if (!jobs || count < 1) return fail(NBH_ERR_ARG, "bad hgcd level batch");
    std::unique_ptr<nbh_hgcd_s, void (*)(nbh_hgcd_s *)> H(new nbh_hgcd_s, hgcd_free);
    size_t words = 0;
    for (int64_t i = 0; i < count; i++) {
        const nbh_hgcd_level_job &q = jobs[i];
        if (q.mn < 1 || q.m1n < 1 || q.p < 1 || q.nn < 1 || !q.m00 || !q.m01 || !q.m10 || !q.m11 || !q.n00 ||
            !q.n01 || !q.n10 || !q.n11 || !q.a || !q.b || !q.ra || !q.rb || !q.sa || !q.sb || !q.n_ab || !q.n_m)
            return fail(NBH_ERR_ARG, "bad hgcd level instance " + std::to_string(i));
        words += 4 * (size_t)(q.p + q.m1n) + 8 * (size_t)(q.mn + q.m1n);
    }
    HIPCHK(hipMalloc(&H->tmp, 8 * words));
    std::vector<nbh_job> prods;
    std::vector<CombineJob> comb;
    uint64_t *t = (uint64_t *)H->tmp;
    // full lengths as bit bounds: the contents may change between runs
    auto job = [&](uint64_t *r, const uint64_t *x, int64_t lx, const uint64_t *y, int64_t ly) {
        prods.push_back(nbh_job{r, x, y, lx, ly, 64 * lx, 64 * ly, 0});
    };
    for (int64_t i = 0; i < count; i++) {
        const nbh_hgcd_level_job &q = jobs[i];
        // hgcd_matrix_adjust!(M1, a, b, p, ra, rb, nn): n11 a, n01 b, n00 b, n10 a
        const int64_t lp = q.p + q.m1n;
        uint64_t *P[4] = {t, t + lp, t + 2 * lp, t + 3 * lp};
        t += 4 * lp;
        job(P[0], q.a, q.p, q.n11, q.m1n);
        job(P[1], q.b, q.p, q.n01, q.m1n);
        job(P[2], q.b, q.p, q.n00, q.m1n);
        job(P[3], q.a, q.p, q.n10, q.m1n);
        const int64_t L = q.p + std::max(q.m1n + 1, q.nn) + 1;
        comb.push_back(CombineJob{P[0], P[1], lp, lp, q.a, L, q.sa, 1, 0, q.ra, q.nn, q.p, q.n_ab});
        comb.push_back(CombineJob{P[2], P[3], lp, lp, q.b, L, q.sb, 1, 0, q.rb, q.nn, q.p, q.n_ab});
        // hgcd_matrix_mul!(M, M1): row (r0, r1) times M1's columns (hgcd_mul2!)
        const int64_t lq = q.mn + q.m1n;
        uint64_t *Q[8];
        for (int k = 0; k < 8; k++) Q[k] = t + k * lq;
        t += 8 * lq;
        uint64_t *row[2][2] = {{q.m00, q.m01}, {q.m10, q.m11}};
        const uint64_t *col[2][2] = {{q.n00, q.n10}, {q.n01, q.n11}};
        for (int rr = 0; rr < 2; rr++)
            for (int cc = 0; cc < 2; cc++) {
                uint64_t *A = Q[4 * rr + 2 * cc], *B = Q[4 * rr + 2 * cc + 1];
                job(A, row[rr][0], q.mn, col[cc][0], q.m1n);
                job(B, row[rr][1], q.mn, col[cc][1], q.m1n);
                comb.push_back(CombineJob{A, B, lq, lq, row[rr][cc], lq + 1, nullptr, 0, 0, nullptr, 0, 0, q.n_m});
            }
    }
    RCCHK(batch_create(prods.data(), (int64_t)prods.size(), &H->products));
    sort_combine_jobs(comb);
    H->ncomb = (int64_t)comb.size();
    H->lens = true;
    HIPCHK(hipMalloc((void **)&H->dcomb, sizeof(CombineJob) * comb.size()));
    HIPCHK(hipMemcpy(H->dcomb, comb.data(), sizeof(CombineJob) * comb.size(), hipMemcpyHostToDevice));
    *out = H.release();
    return NBH_OK;
}

}  // namespace

// ---------------------------------------------------------------------------
// C ABI

extern "C" {

int32_t nbh_init(int32_t device) {
    NBH_GUARD_BEGIN
        std::lock_guard<std::recursive_mutex> lk(g.mu);
        if (g.init) {
            if (device != g.device) return fail(NBH_ERR_ARG, "already initialised on another device");
            return NBH_OK;
        }
        int n = 0;
        HIPCHK(hipGetDeviceCount(&n));
        if (device < 0 || device >= n) return fail(NBH_ERR_ARG, "no such HIP device");
        HIPCHK(hipSetDevice(device));
        if (device < SCRATCH_MAX_DEVICES && !g.scratch_dev[device]) {
            // with `device` current: its copy of the static workspace
            void *ptr = nullptr;
            HIPCHK(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
            if (!ptr) return fail(NBH_ERR_HIP, "no device address for the no-allocation workspace");
            g.scratch_dev[device] = ptr;
        }
        hipDeviceProp_t prop;
        HIPCHK(hipGetDeviceProperties(&prop, device));
        g.device = device;
        g.cus = prop.multiProcessorCount;
        g.lds_block = prop.sharedMemPerBlock;
        g.hbm = prop.totalGlobalMem;
        g.name = prop.gcnArchName;
        if (1.0 / P1_D != P1_INV || 1.0 / P2_D != P2_INV) return fail(NBH_ERR_HIP, "fl(1/p) constant mismatch");
        // A blocking stream: it orders with the legacy default stream (handle 0),
        // which is where callers such as torch enqueue by default.
        HIPCHK(hipStreamCreate(&g.stream));
        HIPCHK(hipEventCreateWithFlags(&g.ev_in, hipEventDisableTiming));
        HIPCHK(hipEventCreateWithFlags(&g.ev_out, hipEventDisableTiming));
        HIPCHK(hipEventCreateWithFlags(&g.last, hipEventDisableTiming));
        g.last_valid = g.last_on_lib = false;
        for (int i = 0; i < NSIDE; i++) HIPCHK(hipStreamCreateWithFlags(&g.side[i], hipStreamNonBlocking));
        HIPCHK(set_lds_limit(g.lds_block));
        load_tuning_file();
        // keep the default tiles inside the device's per-block LDS
        while (8 * (size_t)lds_doubles(g.tun.mid_max) > g.lds_block && g.tun.mid_max > 256) g.tun.mid_max >>= 1;
        while (8 * (size_t)lds_doubles(g.tun.grp_lds) > g.lds_block && g.tun.grp_lds > 256) g.tun.grp_lds >>= 1;
        g.init = true;
        return NBH_OK;
    NBH_GUARD_END
}

int32_t nbh_shutdown(void) {
    std::lock_guard<std::recursive_mutex> lk(g.mu);
    if (!g.init) return NBH_OK;
    if (g.live > 0)
        return fail(NBH_ERR_ARG, std::to_string(g.live) + " live handle(s) still point into the plan cache: "
                                 "destroy them before nbh_shutdown");
    // single products may have run on callers' streams: wait for all of them
    (void)hipDeviceSynchronize();
    for (auto &e : g.plans) (void)hipFree(e.second->dev);
    for (auto &e : g.stages) (void)hipFree(e.second->twc);
    for (auto &e : g.twtab) (void)hipFree(e.second);
    g.plans.clear();
    g.stages.clear();
    g.twtab.clear();
    if (g.ws) (void)hipFree(g.ws);
    if (g.io) (void)hipFree(g.io);
    if (g.pin) (void)hipHostFree(g.pin);
    for (auto &h : g.hs) {
        if (h.pin) (void)hipHostFree(h.pin);
        if (h.dev) (void)hipFree(h.dev);
        if (h.ev) (void)hipEventDestroy(h.ev);
        h = Lib::HostSlot{};
    }
    g.ws = g.io = nullptr;
    g.pin = g.pin_dev = nullptr;
    g.ws_size = g.io_size = g.pin_size = 0;
    for (int i = 0; i < NSIDE; i++) (void)hipStreamDestroy(g.side[i]);
    (void)hipEventDestroy(g.ev_in);
    (void)hipEventDestroy(g.ev_out);
    (void)hipEventDestroy(g.last);
    g.last_valid = false;
    (void)hipStreamDestroy(g.stream);
    g.init = false;
    g.device = -1;
    return NBH_OK;
}

const char *nbh_last_error_string(void) { return t_err.c_str(); }

int32_t nbh_ntt_len(int64_t T, int64_t *N) {
    NBH_GUARD_BEGIN
        if (!N || T < 1) return fail(NBH_ERR_ARG, "bad argument");
        const int64_t n = ntt_len(T);
        if (n < 0) return fail(NBH_ERR_LENGTH, "fp NTT length exceeds the 315*2^34 the primes support");
        *N = n;
        return NBH_OK;
    NBH_GUARD_END
}

int32_t nbh_ntt_params(int64_t bits_a, int64_t bits_b, int32_t *b, int64_t *N) {
    NBH_GUARD_BEGIN
        if (!b || !N || bits_a < 1 || bits_b < 1) return fail(NBH_ERR_ARG, "bad argument");
        int bb = 0;
        RCCHK(ntt_params(bits_a, bits_b, &bb, N));
        *b = bb;
        return NBH_OK;
    NBH_GUARD_END
}

int32_t nbh_mul(uint64_t *r, const uint64_t *a, int64_t m, const uint64_t *b, int64_t n) {
    NBH_GUARD_BEGIN
        std::lock_guard<std::recursive_mutex> lk(g.mu);
        RCCHK(ensure_init());
        return host_product(r, a, m, b, n, false);
    NBH_GUARD_END
}

int32_t nbh_sqr(uint64_t *r, const uint64_t *a, int64_t n) {
    NBH_GUARD_BEGIN
        std::lock_guard<std::recursive_mutex> lk(g.mu);
        RCCHK(ensure_init());
        return host_product(r, a, n, nullptr, n, true);
    NBH_GUARD_END
}

int32_t nbh_mul_dev(uint64_t *r, const uint64_t *a, int64_t m, const uint64_t *b, int64_t n, int64_t bits_a,
                    int64_t bits_b, void *stream) {
    NBH_GUARD_BEGIN
        std::lock_guard<std::recursive_mutex> lk(g.mu);
        RCCHK(ensure_init());
        return run_on(stream, [&](hipStream_t s) { return run_single(r, a, m, b, n, bits_a, bits_b, false, nullptr, 0, s); });
    NBH_GUARD_END
}

int32_t nbh_mul_dev_noalloc(uint64_t *r, const uint64_t *a, int64_t m, const uint64_t *b, int64_t n,
                            int64_t bits_a, int64_t bits_b, void *stream) {
    NBH_GUARD_BEGIN
        std::lock_guard<std::recursive_mutex> lk(g.mu);
        RCCHK(ensure_init());
        void *p = g.device >= 0 && g.device < SCRATCH_MAX_DEVICES ? g.scratch_dev[g.device] : nullptr;
        if (!p) return fail(NBH_ERR_HIP, "the no-allocation workspace has no address on this device");
        const Scratch sc{p, NOALLOC_SCRATCH_BYTES};
        const hipStream_t s = stream ? (hipStream_t)stream : g.stream;
        // ordered after the previous single-product call, like nbh_mul_dev
        // (no std::function here: its capture could take host memory)
        RCCHK(lib_begin(s));
        const int rc = run_single(r, a, m, b, n, bits_a, bits_b, false, nullptr, 0, s, &sc);
        const int rc2 = lib_end(s);
        return rc != NBH_OK ? rc : rc2;
    NBH_GUARD_END
}

int32_t nbh_sqr_dev(uint64_t *r, const uint64_t *a, int64_t n, int64_t bits_a, void *stream) {
    NBH_GUARD_BEGIN
        std::lock_guard<std::recursive_mutex> lk(g.mu);
        RCCHK(ensure_init());
        return run_on(stream,
                      [&](hipStream_t s) { return run_single(r, a, n, nullptr, n, bits_a, bits_a, true, nullptr, 0, s); });
    NBH_GUARD_END
}

int32_t nbh_mullo(uint64_t *r, const uint64_t *a, int64_t la, const uint64_t *b, int64_t lb, int64_t k) {
    NBH_GUARD_BEGIN
        std::lock_guard<std::recursive_mutex> lk(g.mu);
        RCCHK(ensure_init());
        if (k < 1) return fail(NBH_ERR_ARG, "k < 1");
        return host_product(r, a, la, b, lb, false, k);
    NBH_GUARD_END
}

int32_t nbh_sqrlo(uint64_t *r, const uint64_t *a, int64_t la, int64_t k) {
    NBH_GUARD_BEGIN
        std::lock_guard<std::recursive_mutex> lk(g.mu);
        RCCHK(ensure_init());
        if (k < 1) return fail(NBH_ERR_ARG, "k < 1");
        return host_product(r, a, la, nullptr, la, true, k);
    NBH_GUARD_END
}

int32_t nbh_mullo_dev(uint64_t *r, const uint64_t *a, int64_t la, const uint64_t *b, int64_t lb, int64_t k,
                      int64_t bits_a, int64_t bits_b, void *stream) {
    NBH_GUARD_BEGIN
        std::lock_guard<std::recursive_mutex> lk(g.mu);
        RCCHK(ensure_init());
        return run_on(stream, [&](hipStream_t s) { return run_lo(r, a, la, b, lb, k, bits_a, bits_b, false, s); });
    NBH_GUARD_END
}

int32_t nbh_sqrlo_dev(uint64_t *r, const uint64_t *a, int64_t la, int64_t k, int64_t bits_a, void *stream) {
    NBH_GUARD_BEGIN
        std::lock_guard<std::recursive_mutex> lk(g.mu);
        RCCHK(ensure_init());
        return run_on(stream, [&](hipStream_t s) { return run_lo(r, a, la, nullptr, la, k, bits_a, bits_a, true, s); });
    NBH_GUARD_END
}

int32_t nbh_batch_create(const nbh_job *jobs, int64_t count, nbh_batch *out) {
    NBH_GUARD_BEGIN
        std::lock_guard<std::recursive_mutex> lk(g.mu);
        RCCHK(ensure_init());
        if (!out) return fail(NBH_ERR_ARG, "null handle");
        nbh_batch_s *B = nullptr;
        RCCHK(batch_create(jobs, count, &B));
        *out = B;
        g.live++;
        return NBH_OK;
    NBH_GUARD_END
}

int32_t nbh_batch_run(nbh_batch b, void *stream) {
    NBH_GUARD_BEGIN
        std::lock_guard<std::recursive_mutex> lk(g.mu);
        RCCHK(ensure_init());
        if (!b) return fail(NBH_ERR_ARG, "null handle");
        return batch_run(b, stream ? (hipStream_t)stream : g.stream);
    NBH_GUARD_END
}

// Capture the batch's whole launch sequence -- every phase, and the fork /
// join events across the side streams -- into a HIP graph on a private
// stream, so a run is one graph launch.  The sequence records kernels and
// events only (no memset nodes).  Per-launch profiling is off while capturing.
int32_t nbh_batch_capture(nbh_batch b) {
    NBH_GUARD_BEGIN
        std::lock_guard<std::recursive_mutex> lk(g.mu);
        RCCHK(ensure_init());
        if (!b) return fail(NBH_ERR_ARG, "null handle");
        if (b->gexec) return NBH_OK;
        // Restores the profiling flag, ends an open capture and frees the
        // private stream and a partial graph on every way out, an exception
        // included.
        struct CaptureScope {
            hipStream_t cs = nullptr;
            bool capturing = false;
            bool prof = g.prof;
            hipGraph_t graph = nullptr;
            ~CaptureScope() {
                g.prof = prof;
                if (capturing) {
                    hipGraph_t partial = nullptr;
                    if (hipStreamEndCapture(cs, &partial) == hipSuccess && partial) (void)hipGraphDestroy(partial);
                    (void)hipGetLastError();
                }
                if (graph) (void)hipGraphDestroy(graph);
                if (cs) (void)hipStreamDestroy(cs);
            }
        } sc;
        HIPCHK(hipStreamCreateWithFlags(&sc.cs, hipStreamNonBlocking));
        g.prof = false;
        HIPCHK(hipStreamBeginCapture(sc.cs, hipStreamCaptureModeThreadLocal));
        sc.capturing = true;
        const int rc = batch_run(b, sc.cs);
        sc.capturing = false;
        const hipError_t e = hipStreamEndCapture(sc.cs, &sc.graph);
        RCCHK(rc);
        if (e != hipSuccess) return fail(NBH_ERR_HIP, std::string("graph capture: ") + hipGetErrorString(e));
        HIPCHK(hipGraphInstantiate(&b->gexec, sc.graph, nullptr, nullptr, 0));
        b->graph = sc.graph;
        sc.graph = nullptr;  // owned by the batch now
        return NBH_OK;
    NBH_GUARD_END
}

int32_t nbh_batch_replay(nbh_batch b, void *stream) {
    NBH_GUARD_BEGIN
        std::lock_guard<std::recursive_mutex> lk(g.mu);
        RCCHK(ensure_init());
        if (!b) return fail(NBH_ERR_ARG, "null handle");
        if (!b->gexec) RCCHK(nbh_batch_capture(b));
        HIPCHK(hipGraphLaunch(b->gexec, stream ? (hipStream_t)stream : g.stream));
        return NBH_OK;
    NBH_GUARD_END
}

int32_t nbh_batch_arena_bytes(const nbh_job *jobs, int64_t count, int64_t *bytes) {
    NBH_GUARD_BEGIN
        std::lock_guard<std::recursive_mutex> lk(g.mu);
        RCCHK(ensure_init());
        if (!bytes) return fail(NBH_ERR_ARG, "null output");
        size_t need = 0;
        RCCHK(batch_create(jobs, count, nullptr, &need));
        *bytes = (int64_t)need;
        return NBH_OK;
    NBH_GUARD_END
}

int32_t nbh_batch_create_in(const nbh_job *jobs, int64_t count, void *arena, int64_t arena_bytes, void *stream,
                            nbh_batch *out) {
    NBH_GUARD_BEGIN
        std::lock_guard<std::recursive_mutex> lk(g.mu);
        RCCHK(ensure_init());
        if (!out || !arena || arena_bytes <= 0) return fail(NBH_ERR_ARG, "bad arena");
        nbh_batch_s *B = nullptr;
        RCCHK(batch_create(jobs, count, &B, nullptr, arena, (size_t)arena_bytes,
                           stream ? (hipStream_t)stream : g.stream));
        *out = B;
        g.live++;
        return NBH_OK;
    NBH_GUARD_END
}

void nbh_batch_release(nbh_batch b) {
    std::lock_guard<std::recursive_mutex> lk(g.mu);
    if (!b) return;
    batch_free(b);
    g.live--;
}

void nbh_batch_destroy(nbh_batch b) {
    std::lock_guard<std::recursive_mutex> lk(g.mu);
    if (b) {
        (void)hipDeviceSynchronize();
        batch_free(b);
        g.live--;
    }
}

// ---------------------------------------------------------------------------
// Serving queue: continuous batching inside an HBM budget.
//
// Requests (device operands and results) wait in arrival order.  Each poll
// retires the micro-batches whose completion event has fired (their arena
// block is freed and their tickets reported), then, for every idle stream,
// admits the oldest requests while their arena bytes fit the largest free
// block (at most max_batch per micro-batch), plans them inside that block
// (batch_plan's arena form: no device allocation, no host synchronisation)
// and launches them.  So requests that arrive while others run start as soon
// as a stream and memory free up, and micro-batches overlap on the device.

struct nbh_server_s {
    char *arena = nullptr;
    size_t budget = 0;
    int max_batch = 256;
    std::vector<std::pair<size_t, size_t>> free_list;  // (offset, length), sorted, coalesced
    struct Req {
        nbh_job job;
        int64_t ticket;
        size_t need;
    };
    std::deque<Req> pending;
    struct Flight {
        hipStream_t stream = nullptr;
        hipEvent_t ev = nullptr;
        nbh_batch_s *B = nullptr;  // null: the stream is idle
        std::vector<int64_t> tickets;
        size_t off = 0, len = 0;
    };
    std::vector<Flight> flights;
    std::map<std::tuple<int64_t, int64_t, int64_t, int64_t, int64_t, bool>, size_t> need_cache;
    uint64_t cache_gen = 0;    // g.tuning_gen the cached sizes were computed under
    uint64_t pending_gen = 0;  // ... and the waiting requests' sizes
    std::deque<int64_t> done;
    int64_t next_ticket = 0, batches = 0;
};

namespace {

size_t srv_largest(const nbh_server_s &S) {
    size_t m = 0;
    for (const auto &f : S.free_list) m = std::max(m, f.second);
    return m;
}

bool srv_alloc(nbh_server_s &S, size_t n, size_t *off) {
    for (size_t i = 0; i < S.free_list.size(); i++) {
        auto &f = S.free_list[i];
        if (f.second < n) continue;
        *off = f.first;
        if (f.second == n) S.free_list.erase(S.free_list.begin() + (std::ptrdiff_t)i);
        else f = {f.first + n, f.second - n};
        return true;
    }
    return false;
}

void srv_release(nbh_server_s &S, size_t off, size_t n) {
    auto it = std::lower_bound(S.free_list.begin(), S.free_list.end(), std::make_pair(off, (size_t)0));
    it = S.free_list.insert(it, {off, n});
    // merge with the next block, then with the previous one
    if (it + 1 != S.free_list.end() && it->first + it->second == (it + 1)->first) {
        it->second += (it + 1)->second;
        S.free_list.erase(it + 1);
    }
    if (it != S.free_list.begin() && (it - 1)->first + (it - 1)->second == it->first) {
        (it - 1)->second += it->second;
        S.free_list.erase(it);
    }
}

// Arena bytes of one request alone (workspace + tables, 256-aligned); a
// micro-batch needs at most the sum over its requests.
int srv_need(nbh_server_s &S, const nbh_job &j, size_t *need) {
    if (S.cache_gen != g.tuning_gen) {
        // tile counts, and so the tables' bytes, follow the schedule knobs
        S.need_cache.clear();
        S.cache_gen = g.tuning_gen;
    }
    const auto key = std::make_tuple(j.m, j.n, j.bits_a, j.bits_b, j.k, j.b == nullptr);
    auto it = S.need_cache.find(key);
    if (it != S.need_cache.end()) {
        *need = it->second;
        return NBH_OK;
    }
    size_t n = 0;
    RCCHK(batch_create(&j, 1, nullptr, &n));
    n = align256(n);
    S.need_cache[key] = n;
    *need = n;
    return NBH_OK;
}

int srv_retire(nbh_server_s &S) {
    for (auto &f : S.flights) {
        if (!f.B) continue;
        const hipError_t e = hipEventQuery(f.ev);
        if (e == hipErrorNotReady) continue;
        HIPCHK(e);
        for (int64_t t : f.tickets) S.done.push_back(t);
        f.tickets.clear();
        batch_free(f.B);  // host side only: the batch lives in the arena
        f.B = nullptr;
        srv_release(S, f.off, f.len);
    }
    return NBH_OK;
}

// The schedule knobs changed after the waiting requests were sized: size them
// again, so an admitted micro-batch always fits the block it is given.  A
// request that no longer fits the whole budget leaves the queue with an error
// naming its ticket (it would otherwise block every request behind it).
int srv_resize(nbh_server_s &S) {
    if (S.pending_gen == g.tuning_gen) return NBH_OK;
    for (auto it = S.pending.begin(); it != S.pending.end();) {
        size_t need = 0;
        RCCHK(srv_need(S, it->job, &need));
        if (need > S.budget) {
            const int64_t t = it->ticket;
            S.pending.erase(it);
            return fail(NBH_ERR_BUDGET, "ticket " + std::to_string(t) +
                                            " exceeds the server's HBM budget under the new schedule; dropped");
        }
        it->need = need;
        ++it;
    }
    S.pending_gen = g.tuning_gen;
    return NBH_OK;
}

int srv_admit(nbh_server_s &S) {
    RCCHK(srv_resize(S));
    for (auto &f : S.flights) {
        if (f.B || S.pending.empty()) continue;
        const size_t room = srv_largest(S);
        size_t need = 0;
        size_t cnt = 0;
        while (cnt < S.pending.size() && (int)cnt < S.max_batch && need + S.pending[cnt].need <= room)
            need += S.pending[cnt++].need;
        if (cnt == 0) return NBH_OK;  // wait for memory
        size_t off = 0;
        if (!srv_alloc(S, need, &off)) return fail(NBH_ERR_BUDGET, "server arena: no block");
        std::vector<nbh_job> jobs(cnt);
        for (size_t i = 0; i < cnt; i++) jobs[i] = S.pending[i].job;
        nbh_batch_s *B = nullptr;
        const int rc = batch_create(jobs.data(), (int64_t)cnt, &B, nullptr, S.arena + off, need, f.stream);
        if (rc != NBH_OK) {
            srv_release(S, off, need);
            return rc;
        }
        B->nside = 0;  // micro-batches already overlap on the server's streams
        f.B = B;
        f.off = off;
        f.len = need;
        for (size_t i = 0; i < cnt; i++) f.tickets.push_back(S.pending[i].ticket);
        S.pending.erase(S.pending.begin(), S.pending.begin() + (std::ptrdiff_t)cnt);
        RCCHK(batch_run(B, f.stream));
        HIPCHK(hipEventRecord(f.ev, f.stream));
        S.batches++;
    }
    return NBH_OK;
}

void srv_free(nbh_server_s *S) {
    if (!S) return;
    for (auto &f : S->flights) {
        if (f.stream) (void)hipStreamSynchronize(f.stream);
        if (f.B) batch_free(f.B);
        if (f.ev) (void)hipEventDestroy(f.ev);
        if (f.stream) (void)hipStreamDestroy(f.stream);
    }
    if (S->arena) (void)hipFree(S->arena);
    delete S;
}

}  // namespace

int32_t nbh_server_create(int64_t budget_bytes, int32_t nstreams, int32_t max_batch, nbh_server *out) {
    NBH_GUARD_BEGIN
        std::lock_guard<std::recursive_mutex> lk(g.mu);
        RCCHK(ensure_init());
        if (!out || budget_bytes < 4096 || nstreams < 1 || nstreams > 16 || max_batch < 1)
            return fail(NBH_ERR_ARG, "bad server parameters");
        *out = nullptr;
        std::unique_ptr<nbh_server_s, void (*)(nbh_server_s *)> S(new nbh_server_s, srv_free);
        S->budget = (size_t)budget_bytes & ~(size_t)255;
        S->cache_gen = S->pending_gen = g.tuning_gen;
        S->max_batch = max_batch;
        HIPCHK(hipMalloc((void **)&S->arena, S->budget));
        S->free_list.push_back({0, S->budget});
        S->flights.resize((size_t)nstreams);
        for (auto &f : S->flights) {
            HIPCHK(hipStreamCreateWithFlags(&f.stream, hipStreamNonBlocking));
            HIPCHK(hipEventCreateWithFlags(&f.ev, hipEventDisableTiming));
        }
        *out = S.release();
        g.live++;
        return NBH_OK;
    NBH_GUARD_END
}

int32_t nbh_server_submit(nbh_server s, const nbh_job *job, int64_t *ticket) {
    NBH_GUARD_BEGIN
        std::lock_guard<std::recursive_mutex> lk(g.mu);
        RCCHK(ensure_init());
        if (!s || !job || !ticket) return fail(NBH_ERR_ARG, "null argument");
        nbh_job j = *job;
        const bool sqr = j.b == nullptr;
        if (!j.r || !j.a || j.m < 1 || (!sqr && j.n < 1) || j.k < 0) return fail(NBH_ERR_ARG, "bad job");
        if (sqr) j.n = j.m;
        // admission sizes a request from its bit lengths: read the top limbs if not given
        if (j.bits_a <= 0) RCCHK(read_bits(j.a, j.m, &j.bits_a));
        if (sqr) j.bits_b = j.bits_a;
        else if (j.bits_b <= 0) RCCHK(read_bits(j.b, j.n, &j.bits_b));
        // bit lengths are upper bounds: a zero operand is sized as one bit (the
        // engine still computes its zero product exactly)
        j.bits_a = std::max<int64_t>(j.bits_a, 1);
        j.bits_b = std::max<int64_t>(j.bits_b, 1);
        size_t need = 0;
        RCCHK(srv_need(*s, j, &need));
        if (need > s->budget) return fail(NBH_ERR_BUDGET, "request exceeds the server's HBM budget");
        *ticket = s->next_ticket++;
        s->pending.push_back({j, *ticket, need});
        return NBH_OK;
    NBH_GUARD_END
}

int32_t nbh_server_poll(nbh_server s, int64_t *done, int64_t cap, int64_t *ndone) {
    NBH_GUARD_BEGIN
        std::lock_guard<std::recursive_mutex> lk(g.mu);
        RCCHK(ensure_init());
        if (!s || !ndone || cap < 0 || (cap > 0 && !done)) return fail(NBH_ERR_ARG, "bad argument");
        RCCHK(srv_retire(*s));
        RCCHK(srv_admit(*s));
        int64_t n = 0;
        while (n < cap && !s->done.empty()) {
            done[n++] = s->done.front();
            s->done.pop_front();
        }
        *ndone = n;
        return NBH_OK;
    NBH_GUARD_END
}

int32_t nbh_server_drain(nbh_server s) {
    NBH_GUARD_BEGIN
        std::lock_guard<std::recursive_mutex> lk(g.mu);
        RCCHK(ensure_init());
        if (!s) return fail(NBH_ERR_ARG, "null handle");
        for (;;) {
            RCCHK(srv_retire(*s));
            RCCHK(srv_admit(*s));
            bool busy = false;
            for (auto &f : s->flights)
                if (f.B) {
                    busy = true;
                    HIPCHK(hipEventSynchronize(f.ev));
                    break;
                }
            if (!busy && s->pending.empty()) return NBH_OK;
        }
    NBH_GUARD_END
}

int32_t nbh_server_stats(nbh_server s, int64_t *pending, int64_t *in_flight, int64_t *batches) {
    NBH_GUARD_BEGIN
        std::lock_guard<std::recursive_mutex> lk(g.mu);
        if (!s) return fail(NBH_ERR_ARG, "null handle");
        int64_t fl = 0;
        for (auto &f : s->flights) fl += f.B ? (int64_t)f.tickets.size() : 0;
        if (pending) *pending = (int64_t)s->pending.size();
        if (in_flight) *in_flight = fl;
        if (batches) *batches = s->batches;
        return NBH_OK;
    NBH_GUARD_END
}

void nbh_server_destroy(nbh_server s) {
    std::lock_guard<std::recursive_mutex> lk(g.mu);
    if (!s) return;
    srv_free(s);
    g.live--;
}

// Device memory and copies for callers with no GPU array type of their own.
int32_t nbh_alloc(int64_t bytes, void **ptr) {
    NBH_GUARD_BEGIN
        std::lock_guard<std::recursive_mutex> lk(g.mu);
        RCCHK(ensure_init());
        if (!ptr || bytes < 1) return fail(NBH_ERR_ARG, "bad argument");
        *ptr = nullptr;
        if (hipMalloc(ptr, (size_t)bytes) != hipSuccess) {
            (void)hipGetLastError();
            *ptr = nullptr;
            return fail(NBH_ERR_OOM, "device allocation of " + std::to_string(bytes) + " bytes failed");
        }
        return NBH_OK;
    NBH_GUARD_END
}

void nbh_free(void *ptr) {
    std::lock_guard<std::recursive_mutex> lk(g.mu);
    if (ptr) {
        (void)hipDeviceSynchronize();
        (void)hipFree(ptr);
    }
}

int32_t nbh_upload(void *dst, const void *src, int64_t bytes) {
    NBH_GUARD_BEGIN
        std::lock_guard<std::recursive_mutex> lk(g.mu);
        RCCHK(ensure_init());
        if (!dst || !src || bytes < 0) return fail(NBH_ERR_ARG, "bad argument");
        HIPCHK(hipMemcpy(dst, src, (size_t)bytes, hipMemcpyHostToDevice));
        return NBH_OK;
    NBH_GUARD_END
}

int32_t nbh_download(void *dst, const void *src, int64_t bytes) {
    NBH_GUARD_BEGIN
        std::lock_guard<std::recursive_mutex> lk(g.mu);
        RCCHK(ensure_init());
        if (!dst || !src || bytes < 0) return fail(NBH_ERR_ARG, "bad argument");
        HIPCHK(hipDeviceSynchronize());
        HIPCHK(hipMemcpy(dst, src, (size_t)bytes, hipMemcpyDeviceToHost));
        return NBH_OK;
    NBH_GUARD_END
}

int32_t nbh_mul_batch_dev(const nbh_job *jobs, int64_t count, void *stream) {
    NBH_GUARD_BEGIN
        std::lock_guard<std::recursive_mutex> lk(g.mu);
        RCCHK(ensure_init());
        nbh_batch_s *B = nullptr;
        RCCHK(batch_create(jobs, count, &B));
        hipStream_t s = stream ? (hipStream_t)stream : g.stream;
        int rc = batch_run(B, s);
        hipError_t e = hipStreamSynchronize(s);
        batch_free(B);
        if (rc) return rc;
        HIPCHK(e);
        return NBH_OK;
    NBH_GUARD_END
}

// f(j) for j in [0, n) on up to 8 host threads (j interleaved over them) when
// the work is large, else on the calling thread.
static void host_parallel_for(size_t n, bool large, const std::function<void(size_t)> &f) {
    const unsigned hw = std::thread::hardware_concurrency();
    const unsigned nt = large ? std::max(1u, std::min(8u, hw)) : 1u;
    if (nt == 1 || n < 2) {
        for (size_t j = 0; j < n; j++) f(j);
        return;
    }
    std::vector<std::thread> th;
    unsigned started = 1;  // thread 0 is the caller
    try {
        for (unsigned t = 1; t < nt; t++, started++)
            th.emplace_back([&, t] {
                for (size_t j = t; j < n; j += nt) f(j);
            });
    } catch (...) {
        // no thread could be started: the caller takes the remaining residues
    }
    for (unsigned t = 0; t < nt; t++)
        if (t == 0 || t >= started)
            for (size_t j = t; j < n; j += nt) f(j);
    for (auto &x : th) x.join();
}

// Groups of a host-pointer batch (operands + results) and the size from which
// its host staging copies run on several threads.
constexpr size_t HOST_BATCH_GROUP_BYTES = (size_t)128 << 20;
constexpr size_t HOST_PARALLEL_COPY_BYTES = (size_t)8 << 20;

// Host-pointer batch.  Jobs are staged in groups of at most
// HOST_BATCH_GROUP_BYTES (operands + results) through two staging slots: per
// group a host copy of the operands into the slot's pinned buffer, one DMA
// in, one batch planned inside the slot's device buffer (no allocation, no
// host synchronisation), one DMA of the results out and an event; the host
// copies the results out when the slot comes round again, so one group's
// host copies overlap the device work of the group before.  A job larger
// than a group runs on its own with direct copies.  Bit lengths come from the
// operands' top limbs on the host; a product with a zero operand is written
// on the host.
int32_t nbh_mul_batch(const nbh_job *jobs, int64_t count) {
    NBH_GUARD_BEGIN
        std::lock_guard<std::recursive_mutex> lk(g.mu);
        RCCHK(ensure_init());
        if (count < 1 || !jobs) return fail(NBH_ERR_ARG, "empty batch");
        struct Staged {
            nbh_job d;          // sizes and bit lengths; pointers set when the group is placed
            size_t sa, sb, sr;  // aligned bytes of a, b and r in the staging buffer
            int64_t rl;         // result limbs
            bool zero;
        };
        std::vector<Staged> S((size_t)count);
        for (int64_t i = 0; i < count; i++) {
            const nbh_job &q = jobs[i];
            const bool sqr = q.b == nullptr;
            if (!q.r || !q.a || q.m < 1 || (!sqr && q.n < 1) || q.k < 0)
                return fail(NBH_ERR_ARG, "bad job " + std::to_string(i));
            Staged &s = S[(size_t)i];
            // a short product reads only the operands' low k limbs (mullo!, src/mul.jl:308-309)
            const int64_t m = q.k ? std::min(q.m, q.k) : q.m;
            const int64_t n = sqr ? m : (q.k ? std::min(q.n, q.k) : q.n);
            const int64_t ba = bits_from_top(m, q.a[m - 1]);
            const int64_t bb = sqr ? ba : bits_from_top(n, q.b[n - 1]);
            s.d = nbh_job{nullptr, nullptr, nullptr, m, n, ba, bb, q.k};
            s.rl = q.k ? q.k : m + n;
            s.zero = ba == 0 || bb == 0;
            s.sa = s.zero ? 0 : align256(8 * (size_t)m);
            s.sb = s.zero || sqr ? 0 : align256(8 * (size_t)n);
            s.sr = s.zero ? 0 : align256(8 * (size_t)s.rl);
        }
        // A group in flight in a staging slot: its batch, its jobs (device
        // records and caller indices) and where its results sit in the slot.
        struct Pending {
            nbh_batch_s *B = nullptr;
            std::vector<nbh_job> dj;
            std::vector<int64_t> idx;
            char *dev = nullptr;
            bool busy = false, large = false;
        } pend[2];
        // wait for slot k's group, copy its results to the caller, release it
        auto finish = [&](int k) -> int {
            Pending &P = pend[k];
            if (!P.busy) return NBH_OK;
            P.busy = false;
            const hipError_t e = hipEventSynchronize(g.hs[k].ev);
            batch_free(P.B);  // host side only: the batch lives in the slot's arena
            P.B = nullptr;
            HIPCHK(e);
            char *pin = (char *)g.hs[k].pin;
            host_parallel_for(P.dj.size(), P.large, [&](size_t j) {
                std::memcpy(jobs[P.idx[j]].r, pin + ((const char *)P.dj[j].r - P.dev), 8 * (size_t)S[(size_t)P.idx[j]].rl);
            });
            return NBH_OK;
        };
        // on an error: let both slots' work finish before the buffers can be reused
        auto drain = [&]() {
            (void)hipStreamSynchronize(g.stream);
            for (auto &P : pend)
                if (P.busy) {
                    batch_free(P.B);
                    P = Pending{};
                }
        };
        int64_t i0 = 0;
        int slot = 0;
        while (i0 < count) {
            // the group [i0, i1): operands first, then the results, each back to back
            size_t ops = 0, res = 0;
            int64_t i1 = i0;
            while (i1 < count) {
                const Staged &s = S[(size_t)i1];
                if (i1 > i0 && ops + res + s.sa + s.sb + s.sr > HOST_BATCH_GROUP_BYTES) break;
                ops += s.sa + s.sb;
                res += s.sr;
                i1++;
            }
            std::vector<nbh_job> dj;
            std::vector<int64_t> idx;
            for (int64_t i = i0; i < i1; i++) {
                if (S[(size_t)i].zero) {
                    std::memset(jobs[i].r, 0, 8 * (size_t)S[(size_t)i].rl);
                    continue;
                }
                dj.push_back(S[(size_t)i].d);
                idx.push_back(i);
            }
            if (dj.empty()) {
                i0 = i1;
                continue;
            }
            if (ops + res > HOST_BATCH_GROUP_BYTES) {
                // one job larger than a group: on its own, copied directly
                for (int k = 0; k < 2; k++) {
                    const int rc = finish(k);
                    if (rc != NBH_OK) { drain(); return rc; }
                }
                RCCHK(ensure_buf(&g.io, &g.io_size, ops + res));
                char *dev = (char *)g.io;
                const int64_t i = idx[0];
                nbh_job d = dj[0];
                const Staged &st = S[(size_t)i];
                d.a = (const uint64_t *)dev;
                d.b = jobs[i].b ? (const uint64_t *)(dev + st.sa) : nullptr;
                d.r = (uint64_t *)(dev + ops);
                nbh_batch_s *B = nullptr;
                RCCHK(batch_create(&d, 1, &B));
                std::unique_ptr<nbh_batch_s, void (*)(nbh_batch_s *)> hold(B, batch_free);
                HIPCHK(hipMemcpyAsync((void *)d.a, jobs[i].a, 8 * (size_t)d.m, hipMemcpyHostToDevice, g.stream));
                if (jobs[i].b) HIPCHK(hipMemcpyAsync((void *)d.b, jobs[i].b, 8 * (size_t)d.n, hipMemcpyHostToDevice, g.stream));
                int rc = batch_run(B, g.stream);
                if (rc == NBH_OK) {
                    const hipError_t e = hipMemcpyAsync(jobs[i].r, d.r, 8 * (size_t)st.rl, hipMemcpyDeviceToHost, g.stream);
                    if (e != hipSuccess) rc = fail(NBH_ERR_HIP, std::string("copy out: ") + hipGetErrorString(e));
                }
                const hipError_t es = hipStreamSynchronize(g.stream);
                RCCHK(rc);
                HIPCHK(es);
                i0 = i1;
                continue;
            }
            // a staging slot: wait for the group it held two groups ago
            {
                const int rc = finish(slot);
                if (rc != NBH_OK) { drain(); return rc; }
            }
            Lib::HostSlot &H = g.hs[slot];
            // lay the group out: operands, results, then the batch's arena
            std::vector<size_t> offa(dj.size()), offr(dj.size());
            size_t oa = 0, orr = ops;
            for (size_t j = 0; j < dj.size(); j++) {
                const Staged &st = S[(size_t)idx[j]];
                offa[j] = oa;
                offr[j] = orr;
                oa += st.sa + st.sb;
                orr += st.sr;
            }
            auto place = [&](char *base) {
                for (size_t j = 0; j < dj.size(); j++) {
                    dj[j].a = (const uint64_t *)(base + offa[j]);
                    dj[j].b = jobs[idx[j]].b ? (const uint64_t *)(base + offa[j] + S[(size_t)idx[j]].sa) : nullptr;
                    dj[j].r = (uint64_t *)(base + offr[j]);
                }
            };
            size_t need = 0;
            {
                // sizing reads no operand: any non-null base will do
                place(reinterpret_cast<char *>((uintptr_t)4096));
                const int rc = batch_create(dj.data(), (int64_t)dj.size(), nullptr, &need);
                if (rc != NBH_OK) { drain(); return rc; }
            }
            need = align256(need);
            if (H.dev_size < ops + res + need || H.pin_size < ops + res || !H.ev) {
                HIPCHK(hipDeviceSynchronize());
                if (!H.ev) HIPCHK(hipEventCreateWithFlags(&H.ev, hipEventDisableTiming));
                if (H.dev_size < ops + res + need) {
                    if (H.dev) HIPCHK(hipFree(H.dev));
                    H.dev = nullptr;
                    H.dev_size = 0;
                    HIPCHK(hipMalloc(&H.dev, ops + res + need));
                    H.dev_size = ops + res + need;
                }
                if (H.pin_size < ops + res) {
                    if (H.pin) HIPCHK(hipHostFree(H.pin));
                    H.pin = nullptr;
                    H.pin_size = 0;
                    HIPCHK(hipHostMalloc(&H.pin, ops + res, hipHostMallocDefault));
                    H.pin_size = ops + res;
                }
            }
            char *dev = (char *)H.dev;
            place(dev);
            const bool large = ops + res >= HOST_PARALLEL_COPY_BYTES;
            char *pin = (char *)H.pin;
            host_parallel_for(dj.size(), large, [&](size_t j) {
                const nbh_job &q = jobs[idx[j]];
                char *p = pin + ((const char *)dj[j].a - dev);
                std::memcpy(p, q.a, 8 * (size_t)dj[j].m);
                if (q.b) std::memcpy(p + S[(size_t)idx[j]].sa, q.b, 8 * (size_t)dj[j].n);
            });
            // tables into the arena, operands in, the batch, results out, then
            // the slot's event, all in order on g.stream
            nbh_batch_s *B = nullptr;
            int rc = batch_create(dj.data(), (int64_t)dj.size(), &B, nullptr, dev + ops + res, need, g.stream);
            auto hip_ok = [&](hipError_t e, const char *what) {
                if (e != hipSuccess && rc == NBH_OK) rc = fail(NBH_ERR_HIP, std::string(what) + hipGetErrorString(e));
            };
            if (rc == NBH_OK) hip_ok(hipMemcpyAsync(dev, pin, ops, hipMemcpyHostToDevice, g.stream), "copy in: ");
            if (rc == NBH_OK) rc = batch_run(B, g.stream);
            if (rc == NBH_OK) hip_ok(hipMemcpyAsync(pin + ops, dev + ops, res, hipMemcpyDeviceToHost, g.stream), "copy out: ");
            if (rc == NBH_OK) hip_ok(hipEventRecord(H.ev, g.stream), "event: ");
            if (rc != NBH_OK) {
                drain();
                if (B) batch_free(B);
                return rc;
            }
            Pending &P = pend[slot];
            P.B = B;
            P.dj = std::move(dj);
            P.idx = std::move(idx);
            P.dev = dev;
            P.busy = true;
            P.large = large;
            slot ^= 1;
            i0 = i1;
        }
        // the last two groups, oldest first
        for (int k = 0; k < 2; k++) {
            const int rc = finish(slot);
            if (rc != NBH_OK) { drain(); return rc; }
            slot ^= 1;
        }
        return NBH_OK;
    NBH_GUARD_END
}

int32_t nbh_spectrum_create(const uint64_t *b, int64_t n, int64_t bits_b, int64_t bits_a_max, void *stream,
                            nbh_spectrum *out) {
    NBH_GUARD_BEGIN
        std::lock_guard<std::recursive_mutex> lk(g.mu);
        RCCHK(ensure_init());
        if (!out || !b || n < 1 || bits_a_max < 1 || bits_b > 64 * n) return fail(NBH_ERR_ARG, "bad argument");
        *out = nullptr;
        RCCHK(stream_enter(stream));
        if (bits_b < 0) RCCHK(read_bits(b, n, &bits_b));
        std::unique_ptr<nbh_spectrum_s> S(new nbh_spectrum_s);
        S->b = b;
        S->n = n;
        S->bits_b = bits_b;
        S->bits_a_max = bits_a_max;
        if (bits_b == 0) {
            S->zero = 1;
            *out = S.release();
            g.live++;
            return stream_leave(stream);
        }
        int bch = 0;
        int64_t N = 0;
        RCCHK(ntt_params(bits_a_max, bits_b, &bch, &N));
        if (bch < UNPACK_MIN_B) return fail(NBH_ERR_LENGTH, "product too large for device memory (chunk width below 32)");
        S->bch = bch;
        S->N = N;
        S->ncb = (bits_b + bch - 1) / bch;
        PlanEntry *P[2];
        RCCHK(get_plan(N, 0, &P[0]));
        RCCHK(get_plan(N, 1, &P[1]));
        for (int k = 0; k < 2; k++) {
            if (hipMalloc((void **)&S->x[k], sizeof(double) * (size_t)N) != hipSuccess) {
                if (S->x[0]) (void)hipFree(S->x[0]);
                return fail(NBH_ERR_OOM, "spectrum allocation failed");
            }
        }
        // b as operand a of a one-operand job: every forward group, the first
        // one packing the chunks, both primes per launch
        DevJob J;
        std::memset(&J, 0, sizeof J);
        J.a = b;
        J.m = n;
        J.nca = S->ncb;
        J.bch = bch;
        for (int k = 0; k < 2; k++) {
            J.plan[k] = P[k]->dev;
            J.xa[k] = S->x[k];
        }
        const DevPlan &H = P[0]->host;
        for (int gi = 0; gi < H.ngroups; gi++) {
            TileGen tg{gi, 1, H.g[gi].tiles, gi == 0 ? TILE_PACK : 0, 0, J};
            HIPCHK(launch_group(nullptr, nullptr, tg, 2 * H.g[gi].tiles, 1, P[0]->grp_lds_bytes[gi],
                                P[0]->grp_threads[gi], g.stream));
        }
        *out = S.release();
        g.live++;
        return stream_leave(stream);
    NBH_GUARD_END
}

void nbh_spectrum_destroy(nbh_spectrum s) {
    if (!s) return;
    std::lock_guard<std::recursive_mutex> lk(g.mu);
    // queued work on any stream may still read the spectrum
    (void)hipDeviceSynchronize();
    for (int k = 0; k < 2; k++)
        if (s->x[k]) (void)hipFree(s->x[k]);
    delete s;
    g.live--;
}

int32_t nbh_spectrum_info(nbh_spectrum s, int64_t *N, int32_t *b, int64_t *n, int64_t *bits_a_max) {
    NBH_GUARD_BEGIN
        if (!s) return fail(NBH_ERR_ARG, "null handle");
        if (N) *N = s->N;
        if (b) *b = s->bch;
        if (n) *n = s->n;
        if (bits_a_max) *bits_a_max = s->bits_a_max;
        return NBH_OK;
    NBH_GUARD_END
}

int32_t nbh_mul_spec_dev(uint64_t *r, const uint64_t *a, int64_t m, int64_t bits_a, nbh_spectrum s, void *stream) {
    NBH_GUARD_BEGIN
        std::lock_guard<std::recursive_mutex> lk(g.mu);
        RCCHK(ensure_init());
        if (!s) return fail(NBH_ERR_ARG, "null spectrum");
        const nbh_spectrum_s *sp = s;
        return run_on(stream, [&](hipStream_t st) { return run_single(r, a, m, nullptr, 0, bits_a, 0, false, sp, 0, st); });
    NBH_GUARD_END
}

int32_t nbh_spec_batch_create(const nbh_spec_job *jobs, int64_t count, nbh_batch *out) {
    NBH_GUARD_BEGIN
        std::lock_guard<std::recursive_mutex> lk(g.mu);
        RCCHK(ensure_init());
        if (!out || !jobs || count < 1) return fail(NBH_ERR_ARG, "bad argument");
        std::vector<JobIn> in((size_t)count);
        for (int64_t i = 0; i < count; i++) {
            const nbh_spec_job &q = jobs[i];
            if (!q.s) return fail(NBH_ERR_ARG, "job " + std::to_string(i) + " has no spectrum");
            in[(size_t)i] = JobIn{q.r, q.a, q.s->b, q.m, q.s->n, q.bits_a, q.s->bits_b, q.s, 0};
        }
        nbh_batch_s *B = nullptr;
        RCCHK(batch_plan(in.data(), count, &B, nullptr, nullptr, 0, nullptr));
        *out = B;
        g.live++;
        return NBH_OK;
    NBH_GUARD_END
}

// ---------------------------------------------------------------------------
// Barrett reduction (barrett_reduce!, src/barrett.jl:44-81) with the modulus
// and reciprocal held as cached spectra.

struct nbh_barrett_s {
    const uint64_t *m = nullptr;
    int64_t k = 0, lmu = 0;
    nbh_spectrum_s *smu = nullptr, *sm = nullptr;
};

struct nbh_barrett_batch_s {
    nbh_batch_s *p1 = nullptr, *p2 = nullptr;  // q1*mu, q3*m
    void *ws = nullptr;                        // q2 and q3*m per job
    BarrettFixJob *dfix = nullptr;
    int64_t count = 0;
};

namespace {
// Host and device resources of a Barrett batch; the caller holds g.mu and has
// made sure no queued work still uses them.
void barrett_batch_free(nbh_barrett_batch_s *rb) {
    if (!rb) return;
    batch_free(rb->p1);
    batch_free(rb->p2);
    if (rb->ws) (void)hipFree(rb->ws);
    if (rb->dfix) (void)hipFree(rb->dfix);
    delete rb;
}
}  // namespace

int32_t nbh_barrett_create(const uint64_t *m, int64_t k, const uint64_t *mu, int64_t lmu, void *stream,
                           nbh_barrett *out) {
    NBH_GUARD_BEGIN
        std::lock_guard<std::recursive_mutex> lk(g.mu);
        RCCHK(ensure_init());
        if (!out || !m || !mu || k < 1 || lmu < 1 || lmu > k + 2) return fail(NBH_ERR_ARG, "bad argument");
        *out = nullptr;
        std::unique_ptr<nbh_barrett_s> B(new nbh_barrett_s);
        B->m = m;
        B->k = k;
        B->lmu = lmu;
        // q1 = floor(T / beta^(k-1)) has k+1 limbs; q3 = floor(q1 mu / beta^(k+1)) has lmu
        int32_t rc = nbh_spectrum_create(mu, lmu, -1, 64 * (k + 1), stream, &B->smu);
        if (rc == NBH_OK) rc = nbh_spectrum_create(m, k, -1, 64 * lmu, stream, &B->sm);
        if (rc != NBH_OK) {
            nbh_spectrum_destroy(B->smu);
            return rc;
        }
        *out = B.release();
        g.live++;
        return NBH_OK;
    NBH_GUARD_END
}

void nbh_barrett_destroy(nbh_barrett b) {
    if (!b) return;
    std::lock_guard<std::recursive_mutex> lk(g.mu);
    nbh_spectrum_destroy(b->smu);
    nbh_spectrum_destroy(b->sm);
    delete b;
    g.live--;
}

int32_t nbh_barrett_batch_create(nbh_barrett b, const nbh_barrett_job *jobs, int64_t count, nbh_barrett_batch *out) {
    NBH_GUARD_BEGIN
        std::lock_guard<std::recursive_mutex> lk(g.mu);
        RCCHK(ensure_init());
        if (!b || !jobs || count < 1 || !out) return fail(NBH_ERR_ARG, "bad argument");
        *out = nullptr;
        const int64_t k = b->k, lmu = b->lmu;
        const size_t q2_limbs = (size_t)(k + 1 + lmu), p_limbs = (size_t)(lmu + k);
        const size_t per = align256(8 * q2_limbs) + align256(8 * p_limbs);
        std::unique_ptr<nbh_barrett_batch_s, void (*)(nbh_barrett_batch_s *)> R(new nbh_barrett_batch_s,
                                                                               barrett_batch_free);
        R->count = count;
        HIPCHK(hipMalloc(&R->ws, per * (size_t)count));
        std::vector<JobIn> j1((size_t)count), j2((size_t)count);
        std::vector<BarrettFixJob> fx((size_t)count);
        for (int64_t i = 0; i < count; i++) {
            if (!jobs[i].t || !jobs[i].r) return fail(NBH_ERR_ARG, "bad reduction " + std::to_string(i));
            uint64_t *q2 = (uint64_t *)((char *)R->ws + per * (size_t)i);
            uint64_t *p = (uint64_t *)((char *)q2 + align256(8 * q2_limbs));
            // operands are taken at their full lengths: the values change from run
            // to run, and a length is an upper bound of the magnitude
            j1[(size_t)i] = JobIn{q2, jobs[i].t + (k - 1), b->smu->b, k + 1, lmu, 64 * (k + 1), b->smu->bits_b, b->smu, 0};
            // r = T - q3*m is taken mod beta^(k+1): q3*m's low k+1 limbs suffice
            j2[(size_t)i] = JobIn{p, q2 + (k + 1), b->m, lmu, k, 64 * lmu, b->sm->bits_b, b->sm, k + 1};
            fx[(size_t)i] = BarrettFixJob{jobs[i].t, p, p, jobs[i].r, b->m, k};
        }
        RCCHK(batch_plan(j1.data(), count, &R->p1, nullptr, nullptr, 0, nullptr));
        RCCHK(batch_plan(j2.data(), count, &R->p2, nullptr, nullptr, 0, nullptr));
        HIPCHK(hipMalloc((void **)&R->dfix, sizeof(BarrettFixJob) * (size_t)count));
        HIPCHK(hipMemcpy(R->dfix, fx.data(), sizeof(BarrettFixJob) * (size_t)count, hipMemcpyHostToDevice));
        *out = R.release();
        g.live++;
        return NBH_OK;
    NBH_GUARD_END
}

int32_t nbh_barrett_batch_run(nbh_barrett_batch rb, void *stream) {
    NBH_GUARD_BEGIN
        std::lock_guard<std::recursive_mutex> lk(g.mu);
        RCCHK(ensure_init());
        if (!rb) return fail(NBH_ERR_ARG, "null handle");
        hipStream_t s = stream ? (hipStream_t)stream : g.stream;
        RCCHK(batch_run(rb->p1, s));
        RCCHK(batch_run(rb->p2, s));
        HIPCHK(launch_barrett_fix(rb->dfix, rb->count, s));
        return NBH_OK;
    NBH_GUARD_END
}

void nbh_barrett_batch_destroy(nbh_barrett_batch rb) {
    if (!rb) return;
    // shares g.ev_pool with every other batch, and queued work may still use it
    std::lock_guard<std::recursive_mutex> lk(g.mu);
    (void)hipDeviceSynchronize();
    barrett_batch_free(rb);
    g.live--;
}

int32_t nbh_hgcd_create(const nbh_hgcd_job *jobs, int64_t count, int32_t mode, nbh_hgcd *out) {
    NBH_GUARD_BEGIN
        std::lock_guard<std::recursive_mutex> lk(g.mu);
        RCCHK(ensure_init());
        if (!out) return fail(NBH_ERR_ARG, "null handle");
        nbh_hgcd_s *H = nullptr;
        RCCHK(hgcd_create(jobs, count, mode, &H));
        *out = H;
        g.live++;
        return NBH_OK;
    NBH_GUARD_END
}

int32_t nbh_hgcd_run(nbh_hgcd h, void *stream) {
    NBH_GUARD_BEGIN
        std::lock_guard<std::recursive_mutex> lk(g.mu);
        RCCHK(ensure_init());
        if (!h) return fail(NBH_ERR_ARG, "null handle");
        hipStream_t s = stream ? (hipStream_t)stream : g.stream;
        RCCHK(batch_run(h->products, s));
        ProfScope ps(5, s);
        HIPCHK(launch_combine(h->dcomb, h->ncomb, s, h->lens));
        return NBH_OK;
    NBH_GUARD_END
}

int32_t nbh_hgcd_level_create(const nbh_hgcd_level_job *jobs, int64_t count, nbh_hgcd *out) {
    NBH_GUARD_BEGIN
        std::lock_guard<std::recursive_mutex> lk(g.mu);
        RCCHK(ensure_init());
        if (!out) return fail(NBH_ERR_ARG, "null handle");
        nbh_hgcd_s *H = nullptr;
        RCCHK(hgcd_level_create(jobs, count, &H));
        *out = H;
        g.live++;
        return NBH_OK;
    NBH_GUARD_END
}

void nbh_hgcd_destroy(nbh_hgcd h) {
    std::lock_guard<std::recursive_mutex> lk(g.mu);
    if (h) {
        (void)hipDeviceSynchronize();
        hgcd_free(h);
        g.live--;
    }
}

int32_t nbh_ntt_fwd_dev(double *x, int64_t N, int32_t prime, void *stream) {
    NBH_GUARD_BEGIN
        std::lock_guard<std::recursive_mutex> lk(g.mu);
        RCCHK(ensure_init());
        if (!x || (prime != 0 && prime != 1)) return fail(NBH_ERR_ARG, "bad argument");
        PlanEntry *P = nullptr;
        RCCHK(get_plan(N, prime, &P));
        RCCHK(stream_enter(stream));
        DevJob J;
        std::memset(&J, 0, sizeof J);
        J.plan[prime] = P->dev;
        J.xa[prime] = x;
        const DevPlan &H = P->host;
        for (int gi = 0; gi < H.ngroups; gi++) {
            TileGen tg{gi, 1, H.g[gi].tiles, 0, prime, J};
            HIPCHK(launch_group(nullptr, nullptr, tg, H.g[gi].tiles, 1, P->grp_lds_bytes[gi], P->grp_threads[gi],
                                g.stream));
        }
        return stream_leave(stream);
    NBH_GUARD_END
}

int32_t nbh_ntt_rev_dev(double *x, int64_t N, int32_t prime, void *stream) {
    NBH_GUARD_BEGIN
        std::lock_guard<std::recursive_mutex> lk(g.mu);
        RCCHK(ensure_init());
        if (!x || (prime != 0 && prime != 1)) return fail(NBH_ERR_ARG, "bad argument");
        PlanEntry *P = nullptr;
        RCCHK(get_plan(N, prime, &P));
        RCCHK(stream_enter(stream));
        DevJob J;
        std::memset(&J, 0, sizeof J);
        J.plan[prime] = P->dev;
        J.xa[prime] = x;
        const DevPlan &H = P->host;
        for (int gi = H.ngroups - 1; gi >= 0; gi--) {
            TileGen tg{gi, 1, H.g[gi].tiles, 0, prime, J};
            HIPCHK(launch_group(nullptr, nullptr, tg, H.g[gi].tiles, 0, P->grp_lds_bytes[gi], P->grp_threads[gi],
                                g.stream));
        }
        return stream_leave(stream);
    NBH_GUARD_END
}

int32_t nbh_pack_dev(double *x, int64_t N, const uint64_t *limbs, int64_t n, int32_t b, int64_t nch, void *stream) {
    NBH_GUARD_BEGIN
        std::lock_guard<std::recursive_mutex> lk(g.mu);
        RCCHK(ensure_init());
        if (!x || !limbs || n < 1 || b < 1 || b > 52 || nch < 0 || nch > N || (nch - 1) * b >= 64 * n)
            return fail(NBH_ERR_ARG, "bad argument");
        RCCHK(stream_enter(stream));
        HIPCHK(launch_pack(x, N, limbs, n, b, nch, g.stream));
        return stream_leave(stream);
    NBH_GUARD_END
}

int32_t nbh_pointwise_dev(double *xa, const double *xb, int64_t N, int32_t prime, void *stream) {
    NBH_GUARD_BEGIN
        std::lock_guard<std::recursive_mutex> lk(g.mu);
        RCCHK(ensure_init());
        if (!xa || !xb || N < 0 || (prime != 0 && prime != 1)) return fail(NBH_ERR_ARG, "bad argument");
        RCCHK(stream_enter(stream));
        HIPCHK(launch_pointwise(prime, xa, xb, N, g.stream));
        return stream_leave(stream);
    NBH_GUARD_END
}

int32_t nbh_unpack_dev(uint64_t *r, int64_t rn, const double *x1, const double *x2, int64_t N, int64_t nconv,
                       int32_t b, void *stream) {
    NBH_GUARD_BEGIN
        std::lock_guard<std::recursive_mutex> lk(g.mu);
        RCCHK(ensure_init());
        if (!r || !x1 || !x2 || rn < 1 || nconv < 1 || nconv > N || b < UNPACK_MIN_B || b > 49)
            return fail(NBH_ERR_ARG, "bad argument");
        PlanEntry *P[2];
        RCCHK(get_plan(N, 0, &P[0]));
        RCCHK(get_plan(N, 1, &P[1]));
        RCCHK(stream_enter(stream));
        DevJob J;
        std::memset(&J, 0, sizeof J);
        J.r = r;
        J.rn = rn;
        J.nconv = nconv;
        J.bch = b;
        J.bmagic = ~0ull / (uint64_t)b + 1;
        J.nchunks = (rn + UNPACK_CHUNK - 1) / UNPACK_CHUNK;
        J.plan[0] = P[0]->dev;
        J.plan[1] = P[1]->dev;
        J.xa[0] = const_cast<double *>(x1);
        J.xa[1] = const_cast<double *>(x2);
        RCCHK(ensure_buf(&g.ws, &g.ws_size, align256(sizeof(uint32_t) * (size_t)J.nchunks)));
        J.chunk_flags = (uint32_t *)g.ws;
        TileGen ug{0, 1, J.nchunks, 0, 0, J};
        HIPCHK(launch_unpack(nullptr, nullptr, ug, J.nchunks, b, g.stream));
        HIPCHK(launch_carry_fix(nullptr, nullptr, ug, J.nchunks, g.stream));
        return stream_leave(stream);
    NBH_GUARD_END
}

int32_t nbh_plan_nstages(int64_t N, int32_t prime, int32_t *nst) {
    NBH_GUARD_BEGIN
        std::lock_guard<std::recursive_mutex> lk(g.mu);
        RCCHK(ensure_init());
        PlanEntry *P = nullptr;
        RCCHK(get_plan(N, prime, &P));
        *nst = P->host.nst;
        return NBH_OK;
    NBH_GUARD_END
}

int32_t nbh_plan_stage(int64_t N, int32_t prime, int32_t i, int32_t *m, int64_t *span, int64_t *Q, double *rot8) {
    NBH_GUARD_BEGIN
        std::lock_guard<std::recursive_mutex> lk(g.mu);
        RCCHK(ensure_init());
        PlanEntry *P = nullptr;
        RCCHK(get_plan(N, prime, &P));
        if (i < 0 || i >= P->host.nst) return fail(NBH_ERR_ARG, "no such stage");
        const StageDesc &s = P->host.st[i];
        *m = s.m;
        *span = s.L;
        *Q = s.Q;
        std::memcpy(rot8, s.rot, sizeof s.rot);
        return NBH_OK;
    NBH_GUARD_END
}

int32_t nbh_plan_twiddle_row(int64_t N, int32_t prime, int32_t i, int32_t t, double *out) {
    NBH_GUARD_BEGIN
        std::lock_guard<std::recursive_mutex> lk(g.mu);
        RCCHK(ensure_init());
        PlanEntry *P = nullptr;
        RCCHK(get_plan(N, prime, &P));
        if (i < 0 || i >= P->host.nst) return fail(NBH_ERR_ARG, "no such stage");
        const StageDesc &s = P->host.st[i];
        if (t < 1 || t >= s.m || s.m == 2) return fail(NBH_ERR_ARG, "no such twiddle row");
        HIPCHK(hipMemcpy(out, s.tw + (int64_t)(t - 1) * s.Q, sizeof(double) * (size_t)s.Q, hipMemcpyDeviceToHost));
        return NBH_OK;
    NBH_GUARD_END
}

int32_t nbh_plan_ninv(int64_t N, int32_t prime, double *ninv, double *ninvp) {
    NBH_GUARD_BEGIN
        std::lock_guard<std::recursive_mutex> lk(g.mu);
        RCCHK(ensure_init());
        PlanEntry *P = nullptr;
        RCCHK(get_plan(N, prime, &P));
        *ninv = P->host.ninv;
        *ninvp = P->host.ninvp;
        return NBH_OK;
    NBH_GUARD_END
}

int32_t nbh_set_tuning(int32_t grp_lds_doubles, int32_t mid_max, int32_t cmin, int32_t cmax) {
    NBH_GUARD_BEGIN
        std::lock_guard<std::recursive_mutex> lk(g.mu);
        RCCHK(ensure_init());
        const Tuning t{grp_lds_doubles, mid_max, cmin, cmax};
        if (!tuning_ok(t)) return fail(NBH_ERR_ARG, "tuning outside the device limits");
        g.tun = t;
        g.tuning_gen++;
        return NBH_OK;
    NBH_GUARD_END
}

namespace {
int set_table(std::vector<std::pair<int64_t, Tuning>> &dst, const int64_t *nmax, const int32_t *knobs,
              int32_t count) {
    if (count < 0 || (count > 0 && (!nmax || !knobs))) return fail(NBH_ERR_ARG, "bad table");
    std::vector<std::pair<int64_t, Tuning>> tab;
    for (int32_t i = 0; i < count; i++) {
        const Tuning t{knobs[4 * i], knobs[4 * i + 1], knobs[4 * i + 2], knobs[4 * i + 3]};
        if (!tuning_ok(t)) return fail(NBH_ERR_ARG, "table entry outside the device limits");
        if (i > 0 && nmax[i] <= nmax[i - 1]) return fail(NBH_ERR_ARG, "table lengths must increase");
        tab.emplace_back(nmax[i], t);
    }
    dst = tab;
    g.tuning_gen++;
    return NBH_OK;
}

int get_for(int64_t N, int32_t *knobs4, bool single) {
    if (!knobs4) return fail(NBH_ERR_ARG, "null output");
    const Tuning t = tuning_for(N, single);
    knobs4[0] = t.grp_lds;
    knobs4[1] = t.mid_max;
    knobs4[2] = t.cmin;
    knobs4[3] = t.cmax;
    return NBH_OK;
}
}  // namespace

int32_t nbh_set_tuning_table(const int64_t *nmax, const int32_t *knobs, int32_t count) {
    NBH_GUARD_BEGIN
        std::lock_guard<std::recursive_mutex> lk(g.mu);
        RCCHK(ensure_init());
        return set_table(g.table, nmax, knobs, count);
    NBH_GUARD_END
}

int32_t nbh_set_tuning_table_single(const int64_t *nmax, const int32_t *knobs, int32_t count) {
    NBH_GUARD_BEGIN
        std::lock_guard<std::recursive_mutex> lk(g.mu);
        RCCHK(ensure_init());
        return set_table(g.table_single, nmax, knobs, count);
    NBH_GUARD_END
}

int32_t nbh_get_tuning_for(int64_t N, int32_t *knobs4) {
    NBH_GUARD_BEGIN
        std::lock_guard<std::recursive_mutex> lk(g.mu);
        RCCHK(ensure_init());
        return get_for(N, knobs4, false);
    NBH_GUARD_END
}

int32_t nbh_get_tuning_for_single(int64_t N, int32_t *knobs4) {
    NBH_GUARD_BEGIN
        std::lock_guard<std::recursive_mutex> lk(g.mu);
        RCCHK(ensure_init());
        return get_for(N, knobs4, true);
    NBH_GUARD_END
}

int32_t nbh_get_tuning(int32_t *grp_lds_doubles, int32_t *mid_max, int32_t *cmin, int32_t *cmax) {
    NBH_GUARD_BEGIN
        std::lock_guard<std::recursive_mutex> lk(g.mu);
        RCCHK(ensure_init());
        *grp_lds_doubles = g.tun.grp_lds;
        *mid_max = g.tun.mid_max;
        *cmin = g.tun.cmin;
        *cmax = g.tun.cmax;
        return NBH_OK;
    NBH_GUARD_END
}

int32_t nbh_plan_groups(int64_t N, int32_t *ngroups, int64_t *out) {
    NBH_GUARD_BEGIN
        std::lock_guard<std::recursive_mutex> lk(g.mu);
        RCCHK(ensure_init());
        PlanEntry *P = nullptr;
        RCCHK(get_plan(N, 0, &P));
        *ngroups = P->host.ngroups;
        for (int i = 0; i < P->host.ngroups; i++) {
            out[4 * i + 0] = P->host.g[i].La;
            out[4 * i + 1] = P->host.g[i].Lb;
            out[4 * i + 2] = P->host.g[i].S;
            out[4 * i + 3] = P->host.g[i].C;
        }
        return NBH_OK;
    NBH_GUARD_END
}

int32_t nbh_set_streams(int32_t nside) {
    NBH_GUARD_BEGIN
        std::lock_guard<std::recursive_mutex> lk(g.mu);
        RCCHK(ensure_init());
        if (nside < 0 || nside > NSIDE) return fail(NBH_ERR_ARG, "side streams must be 0.." + std::to_string(NSIDE));
        g.nside = nside;
        return NBH_OK;
    NBH_GUARD_END
}

int32_t nbh_batch_set_streams(nbh_batch b, int32_t nside) {
    NBH_GUARD_BEGIN
        std::lock_guard<std::recursive_mutex> lk(g.mu);
        RCCHK(ensure_init());
        if (!b) return fail(NBH_ERR_ARG, "null handle");
        if (nside < -1 || nside > NSIDE) return fail(NBH_ERR_ARG, "side streams must be -1.." + std::to_string(NSIDE));
        b->nside = nside;
        return NBH_OK;
    NBH_GUARD_END
}

int32_t nbh_profile_enable(int32_t on) {
    NBH_GUARD_BEGIN
        std::lock_guard<std::recursive_mutex> lk(g.mu);
        RCCHK(ensure_init());
        g.prof = on != 0;
        return NBH_OK;
    NBH_GUARD_END
}

int32_t nbh_profile_read(double *ms, int64_t *launches, int32_t nkinds) {
    NBH_GUARD_BEGIN
        std::lock_guard<std::recursive_mutex> lk(g.mu);
        RCCHK(ensure_init());
        for (int i = 0; i < nkinds; i++) {
            if (ms) ms[i] = 0;
            if (launches) launches[i] = 0;
        }
        int rc = NBH_OK;
        for (auto &e : g.prof_open) {
            const int kind = std::get<0>(e);
            hipEvent_t a = std::get<1>(e), b = std::get<2>(e);
            if (hipEventSynchronize(b) != hipSuccess) rc = fail(NBH_ERR_HIP, "profile event");
            float t = 0;
            if (rc == NBH_OK && hipEventElapsedTime(&t, a, b) != hipSuccess) rc = fail(NBH_ERR_HIP, "elapsed time");
            if (kind < nkinds) {
                if (ms) ms[kind] += t;
                if (launches) launches[kind] += 1;
            }
            g.prof_pool.push_back(a);
            g.prof_pool.push_back(b);
        }
        g.prof_open.clear();
        return rc;
    NBH_GUARD_END
}

int32_t nbh_synchronize(void *stream) {
    NBH_GUARD_BEGIN
        std::lock_guard<std::recursive_mutex> lk(g.mu);
        RCCHK(ensure_init());
        HIPCHK(hipStreamSynchronize(stream ? (hipStream_t)stream : g.stream));
        return NBH_OK;
    NBH_GUARD_END
}

int32_t nbh_device_info(int32_t *cus, int64_t *lds_per_block, int64_t *hbm_bytes, char *name, int32_t name_len) {
    NBH_GUARD_BEGIN
        std::lock_guard<std::recursive_mutex> lk(g.mu);
        RCCHK(ensure_init());
        if (cus) *cus = g.cus;
        if (lds_per_block) *lds_per_block = (int64_t)g.lds_block;
        if (hbm_bytes) *hbm_bytes = (int64_t)g.hbm;
        if (name && name_len > 0) {
            std::strncpy(name, g.name.c_str(), (size_t)name_len - 1);
            name[name_len - 1] = 0;
        }
        return NBH_OK;
    NBH_GUARD_END
}

}  // extern "C"
