// nbh_internal.h -- device data layout shared by the plan builder, the
// transform/unpack kernels and the C-ABI layer.
//
// A transform of length N = m * 2^k (m | 315) is the reference's flat list of
// DIF stages (src/fpntt.jl:236-271): odd radices 7, 5, 3, 3 at the largest
// spans, then the radix-4 tower, then a twiddle-free radix-2 when k is odd.
// On the GPU the stage list is cut into GROUPS, each one HBM round trip:
// a group's stages are applied to a tile held in LDS.  A group entered at block
// length La and left at block length Lb works on La/Lb "rows" strided Lb apart;
// a tile is C adjacent columns of one such sub-problem, so its global reads are
// C*8-byte contiguous runs.  The last group always has Lb = 1 (contiguous
// blocks) and is fused with the pointwise product and the first inverse group.
#pragma once

#include <stdint.h>

#if defined(__HIPCC__)
#define NBH_HDI __host__ __device__ __forceinline__
#else
#define NBH_HDI inline
#endif

#define NBH_MAX_STAGES 24
#define NBH_MAX_GROUPS 6

namespace nbh {

struct StageDesc {
    int32_t m;          // radix 2, 3, 4, 5 or 7
    int32_t q;          // Q / Lb of the group this stage belongs to (row stride)
    int64_t L;          // span (block length) of the stage
    int64_t Q;          // L / m
    // twiddle row t (1..m-1) is tw[(t-1)*Q + j] = minbal(w_L^(t j)).  The table
    // depends on (prime, m, L) only, so every plan with such a stage shares
    // one copy (the kernels of a batch of many lengths touch one table per
    // span, not one per length).  Null for radix 2.
    const double *tw;
    // offset of the same rows in the plan's contiguous copy DevPlan::twc
    // (lengths the one-workgroup kernel takes): twc[tw_off + (t-1)*Q + j]
    int64_t tw_off;
    uint64_t q_magic;   // ceil(2^32 / q): floor(r / q) = (r * q_magic) >> 32
    double rot[8];      // combine constants (radix-4: rot[0] = wi = r^(N2/4))
};

struct GroupDesc {
    int32_t s0, s1;     // stages [s0, s1) in forward order
    int32_t S;          // rows: La / Lb
    int32_t C;          // columns per tile (power of two dividing Lb)
    int32_t logC;
    int32_t col;        // column group: radices M0 | M1 << 4 | M2 << 8 (0: LDS tiles)
    int64_t La, Lb;
    int64_t tiles;      // tiles per transform: (N / La) * (Lb / C)
};

struct DevPlan {
    int64_t N, N2;
    int32_t prime, nst, ngroups, pad;
    double ninv, ninvp;           // 1/N mod p (minimal balanced) and fl(ninv/p)
    StageDesc st[NBH_MAX_STAGES];
    GroupDesc g[NBH_MAX_GROUPS];
    // every stage's twiddle rows back to back (stage order, radix 2 none),
    // for the one-workgroup kernel's single copy to LDS; null for lengths
    // above SMALL_TWC_MAX_N
    const double *twc;
};

// One product r = a * b (or a^2 when sqr) resident on the device.
struct DevJob {
    const uint64_t *a, *b;
    uint64_t *r;
    int64_t m, n, rn;             // limbs of a, b, r (rn = m + n)
    int64_t nca, ncb, nconv;      // chunk counts and convolution length
    int32_t bch, sqr;             // chunk width in bits; squaring flag
    const DevPlan *plan[2];
    double *xa[2], *xb[2];        // spectra per prime (xb unused for squares)
    uint32_t *chunk_flags;        // unpack chunk carry summaries
    int64_t nchunks;
    int32_t zero;                 // an operand is zero -> r = 0
    int32_t bspec;                // xb holds b's finished forward spectra (shared, read-only)
    uint64_t bmagic;              // ceil(2^64 / bch): floor(x / bch) = umulhi(x, bmagic) for x < 2^57
};

// One unit of work for a launch: a tile of a group of one transform.
struct Tile {
    uint32_t job;
    uint8_t prime, op, group, flags;  // op: 0 operand a, 1 operand b
    uint32_t tile;
    uint32_t pad;
};

// TILE_PACK: the tile's points are packed from limbs.  TILE_CLEAR (a
// single-job mid launch): tile 0 of prime 0 clears the job's chunk flags for
// an unpack that resolves its own carries.
enum TileFlags : uint8_t { TILE_PACK = 1, TILE_CLEAR = 2 };

// Tiles generated from blockIdx when a launch covers a single job: block b is
// (prime, op, tile) = (prime0 + b / (nops*per), (b / per) % nops, b % per).
// A launch over one job also carries that job's record here (its kernels are
// passed no job array), so no device copy of the record has to be stored by
// a launch of its own first.
struct TileGen {
    int32_t group;
    int32_t nops;
    int64_t per;
    int32_t flags;
    int32_t prime0;
    DevJob job;
};

// LDS of one gfx950 CU (bytes).
constexpr size_t LDS_PER_CU = 160 * 1024;

// The mid kernel's block-size bound (its __launch_bounds__): one thread per
// 16 points of a mid tile of at most 16384 points (one LDS array of 8.5
// bytes per point: 139 KB).
constexpr int MID_MAX_THREADS = 1024;

// Block-size bound of the one-workgroup product kernel (k_small), and the
// most stages of a plan it takes (it copies both plans' descriptors to LDS).
constexpr int SMALL_MAX_THREADS = 1024;
constexpr int SMALL_MAX_STAGES = 10;
// Plans of at most this length keep a contiguous copy of their twiddle rows.
constexpr int64_t SMALL_TWC_MAX_N = 4096;

// Doubles of LDS that an n-point tile occupies with one pad double per 16
// (the kernels index LDS as i + i/16).
NBH_HDI constexpr int lds_doubles(int n) { return n + (n >> 4) + 1; }

// Unpack geometry: each thread owns UNPACK_LPT consecutive output limbs.
constexpr int UNPACK_THREADS = 256;
constexpr int UNPACK_LPT = 4;
constexpr int UNPACK_LPC = UNPACK_THREADS * UNPACK_LPT;  // limbs per block-wide pass (basecase, Barrett)
// Threads of a cofactor-step combine block: it walks its output in steps of
// COMBINE_THREADS * UNPACK_LPT limbs, one carry exchange per step (512:
// HGCD steps 3.27 -> 3.24 ms and whole levels 2.97 -> 2.96 ms against 256;
// 1024 was slower on levels).
constexpr int COMBINE_THREADS = 512;

// Chunks per k_carry_fix block (one wave each).
constexpr int CARRY_FIX_WAVES = 4;
// Limbs per unpack chunk: thread 0 of an unpack block takes the 4 limbs just
// below the chunk (the previous chunk's top segment, for its spill only) and
// threads 1..255 the chunk's own limbs.
constexpr int UNPACK_CHUNK = (UNPACK_THREADS - 1) * UNPACK_LPT;
// The smallest chunk width of any product whose spectra fit in device memory.
constexpr int UNPACK_MIN_B = 32;
// Staged coefficients of one unpack chunk at chunk width b (>= the count any
// chunk starts, 64 (UNPACK_CHUNK + 4) / b + 1); the launch's LDS holds two
// arrays of this many words for the smallest b among its jobs.
NBH_HDI constexpr int unpack_stage(int b) { return (UNPACK_CHUNK + UNPACK_LPT) * 64 / b + 2; }
NBH_HDI constexpr size_t unpack_lds_bytes(int bmin) { return 2 * sizeof(uint64_t) * (size_t)unpack_stage(bmin); }
constexpr int GU = 8;  // coefficients per thread per staging round

constexpr uint32_t CHUNK_G = 1;  // carry out with carry-in 0
constexpr uint32_t CHUNK_P = 2;  // every limb all-ones: carry passes through
// Single-job unpack grids up to this many chunks resolve their carries in the
// unpack itself (k_unpack<true>): far fewer blocks than one launch holds
// resident at once (at least four unpack blocks per CU fit, 256 CUs).
constexpr int64_t UNPACK_LB_MAX_CHUNKS = 256;

}  // namespace nbh

namespace nbh {
// A run of consecutive tiles of one (job, prime, operand, group): the host
// plans batches as ranges; a device kernel expands them into the tile list.
struct TileRange {
    Tile head;       // head.tile is the first tile index
    uint32_t count;
    uint32_t out;    // position of the first tile in the expanded list
};

// One output of a 2x2 cofactor step: r = A + B, or r = |A - B| with its sign
// (A, B are products in the step's workspace, zero above their lengths).
struct CombineJob {
    const uint64_t *A, *B;
    int64_t la, lb;
    uint64_t *r;
    int64_t lr;
    int32_t *sign;  // sub only: -1, 0, +1 (device)
    int32_t sub, pad;
    const uint64_t *C;  // when set: A is first replaced by A + C * beta^cshift, written to r
    int64_t lc, cshift;
    // when set: max(*len, 1 + index of r's highest nonzero limb) is stored here
    // (normlen; the combine's init launch first sets every *len to 1)
    int64_t *len;
};

// Barrett correction of one reduction (barrett_reduce!, src/barrett.jl:
// 67-80): w = (T - P) mod beta^(k+1) over the low k+1 limbs, then w -= m while
// w >= m (at most twice), r = w's low k limbs.  w may be P's own buffer.
struct BarrettFixJob {
    const uint64_t *T;  // the reduced value's low k+1 limbs
    const uint64_t *P;  // q3*m, at least k+1 limbs
    uint64_t *w;        // k+1 limbs of working space
    uint64_t *r;        // k limbs out
    const uint64_t *m;  // k limbs
    int64_t k;
};

// A product below the NTT thresholds (mul!: smaller operand < 64 limbs or
// m + n < 304; sqr!: n < 160, src/mul.jl:24-26,127-164,383-389): schoolbook
// on the device.  b is the shorter operand (n <= BASE_NMAX limbs).
constexpr int BASE_NMAX = 160;
constexpr int64_t MUL_FPNTT_MIN = 64;        // src/mul.jl:24
constexpr int64_t MUL_FPNTT_THRESHOLD = 152; // src/mul.jl:25
constexpr int64_t SQR_FPNTT_THRESHOLD = 160; // src/mul.jl:26
struct BaseJob {
    const uint64_t *a, *b;
    uint64_t *r;
    int64_t m, n;  // limbs of a (the longer) and b
    int64_t rn;    // limbs of r written: the product mod beta^rn (m + n for the full product)
};
}  // namespace nbh

static_assert(sizeof(nbh::Tile) == 16, "kernels read a Tile as four 32-bit words");
