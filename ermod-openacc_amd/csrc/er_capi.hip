// er_capi.hip -- host side of the C ABI (include/er_capi.h): context, device
// buffers, per-frame orchestration of the kernels in er_kernels.h.
#include "er_capi.h"
#include "er_kernels.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <string>
#include <type_traits>
#include <vector>

namespace {

constexpr int NI_B = 16;  // insertion interleave of the batched path
constexpr int ER_KSLICES = 16;  // insertion slices of the split-K edens/ecorr sums
constexpr int ER_SPARSE_SITES = 12;  // above this many sites the plane kernel spreads + transforms

// device buffer owned by one context (freed with it); movable, not copyable
template <typename T> struct DBuf {
    T *p = nullptr;
    size_t n = 0;
    DBuf() = default;
    DBuf(const DBuf &) = delete;
    DBuf &operator=(const DBuf &) = delete;
    DBuf(DBuf &&o) noexcept : p(o.p), n(o.n)
    {
        o.p = nullptr;
        o.n = 0;
    }
    DBuf &operator=(DBuf &&o) noexcept
    {
        if (this != &o) {
            release();
            p = o.p;
            n = o.n;
            o.p = nullptr;
            o.n = 0;
        }
        return *this;
    }
    ~DBuf() { release(); }
    hipError_t ensure(size_t count)
    {
        if (count <= n && p) return hipSuccess;
        if (p) (void)hipFree(p);
        p = nullptr;
        n = 0;
        hipError_t e = hipMalloc((void **)&p, std::max<size_t>(count, 1) * sizeof(T));
        if (e == hipSuccess) n = count;
        return e;
    }
    void release()
    {
        if (p) (void)hipFree(p);
        p = nullptr;
        n = 0;
    }
};

}  // namespace

// Scratch of er_run_noalloc: the split-K partial sums of edens/ecorr and the
// weight kernel's block partials, in device memory that the library never
// allocates at run time.  Each GPU holds its own copy for the life of the
// process.  In host code `scratch` names a host-side stand-in, not a device
// address: kernels receive the per-device address that hipGetSymbolAddress
// returns at er_create.  Every er_run_noalloc on one GPU shares the array,
// so such calls must not overlap on that GPU (one stream at a time).
#define ER_NOALLOC_SCRATCH_BYTES ((size_t)64 << 20)
static __device__ __align__(256) double scratch[ER_NOALLOC_SCRATCH_BYTES / sizeof(double)];

namespace {
constexpr int ER_MAX_DEVICES = 64;
std::mutex g_scratch_mu;
double *g_scratch_ptr[ER_MAX_DEVICES] = {};
}  // namespace

// One solute analysed in the context: its sites, its insertions of the
// current frame, its per-insertion energies and its block accumulators.  A
// context has one slot unless er_set_solutes gives it several test
// particles; the frame-level state is shared by all slots.
struct Slot {
    int nslt = 0;                  // sites per insertion / per solute molecule
    float qslt = 0.f;              // refs test-particle charge
    DBuf<float> d_sltq;
    DBuf<int> d_sltlj;
    int nins = 0;
    bool have_ins = false, wslt_given = false;
    DBuf<float> d_sltpos_own, d_wslt;
    const float *d_sltpos = nullptr;
    // uvengy / uvengy0 of the last run: this solute's rows of the run's
    // energies (valid until the next run)
    const float *U_view = nullptr, *u0_view = nullptr;
    int last_NI = 16, last_nins = 0;
    // accumulators
    DBuf<double> d_edens, d_ecorr, d_eself, d_state, d_slnuv;
    DBuf<int> d_flags;
    DBuf<float> d_fstate;
    DBuf<unsigned int> d_minmax;
    // halt key, 'mbin' warning key; captured (key, energy) of each
    DBuf<unsigned long long> d_errkey, d_warnkey, d_halt;
    DBuf<unsigned int> d_seq;      // runs of this slot (device counter: graph replays count too)
    int pair_split = 0;            // pair-search workgroups per insertion (0: built-in choice)
    void release()
    {
        d_sltq.release(); d_sltlj.release(); d_sltpos_own.release(); d_wslt.release(); d_seq.release();
        d_edens.release(); d_ecorr.release(); d_eself.release(); d_state.release();
        d_slnuv.release(); d_flags.release(); d_warnkey.release(); d_fstate.release(); d_minmax.release();
        d_errkey.release(); d_halt.release();
    }
};

struct er_ctx {
    // sizes and parameters
    int numatm = 0, nummol = 0, ntype = 0, numslv = 0;
    er_params prm{};
    int K1 = 0, K2 = 0, K3 = 0, H = 0, ord = 0;
    size_t G = 0, Gc = 0;
    PairParams pp{};
    std::vector<int> molbeg;       // 0-based [nummol+1]
    std::vector<int> uvspec;       // [nummol]
    std::vector<float> molq;       // [nummol]
    std::vector<int> sp_begin, sp_end;
    float qtot = 0.f;              // sum of physical molecule charges

    // device topology
    DBuf<float> d_charge, d_ljene, d_ljlensq, d_molq;
    DBuf<int> d_ljtype, d_molbeg, d_spb, d_spe;

    // solutes (one slot per test particle; slot 0 in soln runs)
    std::vector<Slot> slots = std::vector<Slot>(1);
    int cur = 0;
    Slot &S() { return slots[cur]; }
    int halt_slot = 0;
    int nsltmol = 0;               // soln: number of solute molecules
    std::vector<int> tagslt;       // 0-based
    DBuf<int> d_tag;

    // mesh
    int ermax = 0, esmax = 0;
    std::vector<int> uvmax, uvsoft, idmin;
    std::vector<double> uvcrd, escrd;
    DBuf<double> d_uvcrd, d_escrd;
    DBuf<int> d_uvmax, d_uvsoft, d_idmin;
    bool have_mesh = false;

    // frame
    bool have_frame = false, frame_dirty = false, cell_changed = true;
    float cell[9] = {0}, prev_cell[9] = {0};
    bool have_prev_cell = false;
    FrameConst fc{};
    double invcl_d[9] = {0}, volume_d = 0.0;
    float w_sys = 1.f;
    DBuf<float> d_sitepos_own;
    const float *d_sitepos = nullptr;
    DBuf<float4> d_posn, d_molref;
    DBuf<uint4> d_posu;            // lattice coordinates (pair energies)
    DBuf<float4> d_segref;         // segment reference + radius at each segment's first atom
    DBuf<float> d_splw;
    DBuf<int4> d_splg;
    std::vector<float> h_pin;      // host staging (pageable)

    // insertions
    DBuf<float4> d_insposn, d_insref;
    DBuf<uint4> d_insposu;
    DBuf<float> d_insw, d_selfreal;
    DBuf<int4> d_insg;

    // PME
    std::vector<double> splfc[3];
    DBuf<double> d_splfc1, d_splfc2, d_splfc3;
    DBuf<float> d_engfac;
    DBuf<float2> d_tw1, d_tw2, d_tw3, d_twm, d_tws;
    DBuf<double> d_selfpart;
    DBuf<float2> d_Ftab;           // per-insertion separable structure-factor factors
    DBuf<float2> d_T;
    DBuf<float> d_phi;
    int chunk = 1024;              // insertions per launch sequence (ER_CHUNK overrides)
    int ecorr_kmin = 64;           // least insertions per split-K slice of edens/ecorr (ER_ECORR_KMIN overrides)

    // spatial order / cell list (per frame)
    int nc = 1, ncell = 1;
    float cell_w = 1.f;            // smallest cell width (perpendicular), Angstrom
    float cell_min = 0.f;          // tuned least cell width, Angstrom (0: the built-in 3.5)
    double cell_wmin = 0.0;        // the frame cell's smallest perpendicular width
    DBuf<int> d_perm;
    DBuf<unsigned int> d_keyrank, d_ccnt, d_cstart;
    DBuf<float> d_rmolmax;

    // atoms sorted by PME brick (per frame) for the brick gather
    int nbx = 1, nby = 1, nbz = 1, nbrick = 1;
    int ncu = 0;  // compute units of the context's device
    DBuf<int> d_border;           // bricks in Morton order
    DBuf<uint4> d_bdesc;          // per frame: (brick, first atom, end atom) in Morton order
    DBuf<float> d_brec;           // atom records in brick order
    DBuf<unsigned int> d_akeyrank, d_bcnt, d_bstart;
    DBuf<float> d_Eat;

    // MFMA accumulator layout (lane, register) -> (row, col), measured once
    bool use_mfma = false;
    DBuf<int> d_dmap;

    // energies / histograms
    DBuf<float> d_meshf, d_resself;
    DBuf<double> d_wpart;  // k_weights block partials
    DBuf<unsigned short> d_blut;  // bin brackets per species (k_hist)
    DBuf<double> d_flpart;
    int nslice = 1;
    int hist_slices = 0;           // tuned histogram slices of a refs run (0: built-in choice)
    DBuf<uint32_t> d_C;
    DBuf<double> d_w, d_flceng;

    // accumulators
    DBuf<double> d_part, d_epart;  // split-K partial sums of ecorr / edens

    // the run in progress: its solutes end to end (RunPlan), their device
    // descriptors, the insertions' coordinates and the run's energies
    struct RunPlan {
        int NI = NI_B;
        std::vector<int> order;    // slot ids in run order (few-site solutes first)
        std::vector<SlotDesc> desc;
        int ntot = 0, nsite = 0, ndense = 0, pair_blocks = 0, nz = 0, wrow = 1;
    } plan;
    DBuf<SlotDesc> d_sd;
    std::vector<SlotDesc> sd_up;   // what d_sd holds
    DBuf<float> d_sltcat;          // [nsite][3] the run's insertion coordinates
    DBuf<float> d_Ucat, d_u0cat;   // [ntot/NI][nummol][NI], [ntot]
    bool capturing = false;        // a HIP graph is being recorded
    bool capture_bad = false;      // a call failed while recording
    bool capture_engfac = false;   // the recording computes the influence function before its first run
    int capture_runs = 0;          // runs recorded so far
    bool prof_before_capture = false;
    bool cap_have_prev = false;    // influence-function cell when recording began
    float cap_prev_cell[9] = {0};
    unsigned long long cap_gen = 0;
    // device buffers of the run path are reallocated or freed only where
    // this counter moves; a graph replays only while it is unchanged
    unsigned long long alloc_gen = 0;
    struct GraphRec;
    std::vector<GraphRec *> graphs;
    // er_run_noalloc: this device's address of the file-scope scratch array,
    // and whether the run in progress may allocate
    double *scratch_dev = nullptr;
    bool noalloc = false;

    // diagnostics: event pairs recorded around launches while profiling
    bool prof = false;
    struct ProfRec { int kind; hipEvent_t a, b; };
    std::vector<ProfRec> prof_recs;
    std::vector<hipEvent_t> prof_pool;
    double prof_ms[ER_K_COUNT] = {0};
    int64_t prof_n[ER_K_COUNT] = {0};

    hipStream_t stream = nullptr;
    // the frame's atom preparation (splines, cell list, brick records) runs
    // on a second stream, default priority, beside the insertions' plane
    // DFTs; the main stream waits on ev_atoms before the gather or the pairs
    hipStream_t stream2 = nullptr;
    hipEvent_t ev_in = nullptr, ev_atoms = nullptr;
    bool atoms_pending = false;
    std::string errmsg;
    char halt_code[8] = {0};
    int halt_species = 0;
    float halt_energy = 0.f;
};

static int fail(er_ctx *c, int code, const char *fmt, ...)
{
    if (c) {
        if (c->capturing) c->capture_bad = true;
        char buf[512];
        va_list ap;
        va_start(ap, fmt);
        vsnprintf(buf, sizeof buf, fmt, ap);
        va_end(ap);
        c->errmsg = buf;
    }
    return code;
}

static hipEvent_t prof_event(er_ctx *c)
{
    if (!c->prof_pool.empty()) {
        hipEvent_t e = c->prof_pool.back();
        c->prof_pool.pop_back();
        return e;
    }
    hipEvent_t e = nullptr;
    if (hipEventCreate(&e) != hipSuccess) return nullptr;
    return e;
}

// RAII bracket: records an event pair around the launches in its scope
struct ProfScope {
    er_ctx *c;
    int kind;
    hipEvent_t a = nullptr;
    ProfScope(er_ctx *c_, int k) : c(c_), kind(k)
    {
        if (c->prof && (a = prof_event(c))) (void)hipEventRecord(a, c->stream);
    }
    ~ProfScope()
    {
        if (!a) return;
        hipEvent_t b = prof_event(c);
        if (!b) return;
        (void)hipEventRecord(b, c->stream);
        c->prof_recs.push_back({kind, a, b});
    }
};

static void prof_collect(er_ctx *c)
{
    if (c->prof_recs.empty()) return;
    (void)hipStreamSynchronize(c->stream);
    for (auto &r : c->prof_recs) {
        float ms = 0.f;
        if (hipEventElapsedTime(&ms, r.a, r.b) == hipSuccess) {
            c->prof_ms[r.kind] += ms;
            c->prof_n[r.kind] += 1;
        }
        c->prof_pool.push_back(r.a);
        c->prof_pool.push_back(r.b);
    }
    c->prof_recs.clear();
}

#define HIPCHK(c, x)                                                                         \
    do {                                                                                     \
        hipError_t e_ = (x);                                                                 \
        if (e_ != hipSuccess)                                                                \
            return fail((c), ER_E_HIP, "%s failed: %s (%s:%d)", #x, hipGetErrorString(e_), \
                        __FILE__, __LINE__);                                                 \
    } while (0)

// the entry points that cannot be recorded into a graph
#define ER_NO_CAPTURE(c, who)                                                                          \
    do {                                                                                               \
        if ((c)->capturing)                                                                            \
            return fail((c), ER_E_STATE, "%s: not allowed while a graph is recorded (er_graph_begin)", \
                        who);                                                                          \
    } while (0)

static bool fft_len_ok(int n)
{
    static const int ok[] = {32, 36, 40, 48, 64, 72, 80, 96, 108, 128};
    for (int v : ok)
        if (v == n) return true;
    return false;
}

// (K2, K3) plane shapes with a compiled plane kernel
static bool plane_ok(int k2, int k3) { return (k2 == k3 && fft_len_ok(k2)) || (k2 == 128 && k3 == 108); }

// cardinal B-spline M_n (spline.F90:46-58) in double for the |b|^2 factors
static double mn_spline(int n, double x)
{
    if (!(x > 0.0 && x < (double)n)) return 0.0;
    if (n == 2) return 1.0 - std::fabs(x - 1.0);
    return (x * mn_spline(n - 1, x) + ((double)n - x) * mn_spline(n - 1, x - 1.0)) / (double)(n - 1);
}

// init_spline_axis (recpcal.F90:97-119)
static std::vector<double> spline_axis(int order, int K)
{
    std::vector<double> s(K);
    for (int rci = 0; rci < K; ++rci) {
        double re = 0.0, im = 0.0;
        for (int spi = 0; spi <= order - 2; ++spi) {
            const double chr = mn_spline(order, (double)(spi + 1));
            const double a = 2.0 * M_PI * (double)(spi * rci) / (double)K;
            re += chr * std::cos(a);
            im += chr * std::sin(a);
        }
        s[rci] = re * re + im * im;
    }
    return s;
}

static hipError_t upload_tw(DBuf<float2> &b, int N, hipStream_t s)
{
    std::vector<float2> t(N);
    for (int k = 0; k < N; ++k) {
        const double a = 2.0 * M_PI * (double)k / (double)N;
        t[k] = make_float2((float)std::cos(a), (float)std::sin(a));
    }
    hipError_t e = b.ensure(N);
    if (e != hipSuccess) return e;
    (void)s;
    return hipMemcpy(b.p, t.data(), N * sizeof(float2), hipMemcpyHostToDevice);
}

// GROMACS force-switch constants (realcal.F90:899-915)
static void gmx_params(int pw, float lw, float up, float &A, float &B, float &C)
{
    const float df = up - lw;
    A = -(float)pw * ((float)(pw + 4) * up - (float)(pw + 1) * lw) / (std::pow(up, (float)(pw + 2)) * df * df) / 3.0f;
    B = (float)pw * ((float)(pw + 3) * up - (float)(pw + 1) * lw) / (std::pow(up, (float)(pw + 2)) * df * df * df) / 4.0f;
    C = 1.0f / std::pow(up, (float)pw) - A * df * df * df - B * df * df * df * df;
}

// mesh-independent per-slot state: weights' state (voffset survives engclear),
// flags, halt key and minuv/maxuv
static int slot_state_init(er_ctx *c, Slot &sl)
{
    hipStream_t s = c->stream;
    HIPCHK(c, sl.d_state.ensure(4));
    HIPCHK(c, sl.d_flags.ensure(4));
    HIPCHK(c, sl.d_fstate.ensure(4));
    HIPCHK(c, sl.d_warnkey.ensure(1));
    HIPCHK(c, sl.d_errkey.ensure(1));
    HIPCHK(c, sl.d_halt.ensure(4));
    HIPCHK(c, sl.d_seq.ensure(1));
    HIPCHK(c, sl.d_minmax.ensure(2 * (size_t)(c->numslv + 1)));
    HIPCHK(c, sl.d_slnuv.ensure(c->numslv));
    HIPCHK(c, hipMemsetAsync(sl.d_state.p, 0, 4 * sizeof(double), s));
    HIPCHK(c, hipMemsetAsync(sl.d_flags.p, 0, 4 * sizeof(int), s));
    HIPCHK(c, hipMemsetAsync(sl.d_fstate.p, 0, 4 * sizeof(float), s));
    HIPCHK(c, hipMemsetAsync(sl.d_warnkey.p, 0xff, sizeof(unsigned long long), s));
    HIPCHK(c, hipMemsetAsync(sl.d_errkey.p, 0xff, sizeof(unsigned long long), s));
    HIPCHK(c, hipMemsetAsync(sl.d_halt.p, 0xff, 4 * sizeof(unsigned long long), s));
    HIPCHK(c, hipMemsetAsync(sl.d_seq.p, 0, sizeof(unsigned int), s));
    HIPCHK(c, hipMemsetAsync(sl.d_slnuv.p, 0, c->numslv * sizeof(double), s));
    std::vector<unsigned int> mm(2 * (size_t)(c->numslv + 1));
    auto ord = [](float f) {
        unsigned int u;
        std::memcpy(&u, &f, 4);
        return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
    };
    for (int k = 0; k <= c->numslv; ++k) {
        mm[2 * k] = ord(3.402823466e38f);       // minuv = huge (engproc.F90:299-302)
        mm[2 * k + 1] = ord(-3.402823466e38f);  // maxuv = -huge
    }
    HIPCHK(c, hipMemcpy(sl.d_minmax.p, mm.data(), mm.size() * sizeof(unsigned int), hipMemcpyHostToDevice));
    return ER_OK;
}

// mesh-sized per-slot accumulators (contents cleared by er_clear)
static int slot_acc_alloc(er_ctx *c, Slot &sl)
{
    HIPCHK(c, sl.d_edens.ensure(c->ermax));
    HIPCHK(c, sl.d_ecorr.ensure((size_t)c->ermax * c->ermax));
    HIPCHK(c, sl.d_eself.ensure(std::max(1, c->esmax)));
    return ER_OK;
}

static int slot_clear(er_ctx *c, Slot &sl)
{
    hipStream_t s = c->stream;
    HIPCHK(c, hipMemsetAsync(sl.d_edens.p, 0, c->ermax * sizeof(double), s));
    HIPCHK(c, hipMemsetAsync(sl.d_ecorr.p, 0, (size_t)c->ermax * c->ermax * sizeof(double), s));
    HIPCHK(c, hipMemsetAsync(sl.d_eself.p, 0, std::max(1, c->esmax) * sizeof(double), s));
    HIPCHK(c, hipMemsetAsync(sl.d_slnuv.p, 0, c->numslv * sizeof(double), s));
    // engnorm, engsmpl, avslf (voffset survives, engproc.F90:336-338)
    HIPCHK(c, hipMemsetAsync(sl.d_state.p, 0, 3 * sizeof(double), s));
    return ER_OK;
}

extern "C" int er_create(const er_topology *topo, const er_params *prm, er_ctx **out)
{
    if (!topo || !prm || !out) return ER_E_ARG;
    *out = nullptr;
    er_ctx *c = new er_ctx();
    auto bad = [&](int code) { *out = c; return code; };
    if (topo->numatm <= 0 || topo->nummol <= 0 || topo->ljtype_max <= 0 || topo->numslv <= 0 ||
        !topo->mol_begin_index || !topo->ljtype || !topo->uvspec || !topo->charge || !topo->ljene_mat ||
        !topo->ljlensq_mat)
        return bad(fail(c, ER_E_ARG, "er_create: incomplete topology"));
    if (prm->cltype != 0 && prm->cltype != 2 && prm->cltype != 3)
        return bad(fail(c, ER_E_UNSUPPORTED, "er_create: cltype 0 (bare Coulomb), 2 (PME) and 3 (PPPM) are implemented"));
    if (prm->cltype == 3 && prm->splodr % 2 == 1)
        return bad(fail(c, ER_E_ARG, "er_create: interpolation order for PPPM must be even"));
    if (prm->boxshp == 0) {  // an open system: bare Coulomb at constant volume (setconf.F90:361-364, set_prs)
        if (prm->cltype != 0 || prm->estype == 2)
            return bad(fail(c, ER_E_ARG, "er_create: a non-periodic system needs cltype=0 and estype=1 (set_prs)"));
    } else if (prm->boxshp != 1) {
        return bad(fail(c, ER_E_ARG, "er_create: boxshp must be 0 or 1"));
    }
    if (prm->slttype < 1 || prm->slttype > 3) return bad(fail(c, ER_E_ARG, "er_create: slttype"));
    if (prm->ljswitch < 0 || prm->ljswitch > 3) return bad(fail(c, ER_E_ARG, "er_create: ljswitch"));
    if (prm->splodr < 3 || prm->splodr > 10)
        return bad(fail(c, ER_E_UNSUPPORTED, "er_create: splodr must lie in 3..10"));
    // any grid: the compiled shapes run the four-step kernels, the others the
    // direct-DFT ones, whose plane must fit a workgroup's LDS
    for (int d = 0; d < 3; ++d)
        if (prm->ms[d] < std::max(16, 2 * prm->splodr) || prm->ms[d] > 512)
            return bad(fail(c, ER_E_UNSUPPORTED, "er_create: ms%dmax = %d outside %d..512", d + 1, prm->ms[d],
                            std::max(16, 2 * prm->splodr)));
    if (!plane_ok(prm->ms[1], prm->ms[2]) &&
        (prm->ms[1] > 256 || prm->ms[2] > 256 || spec2d_gen_lds(prm->ms[1], prm->ms[2]) > 160 * 1024))
        return bad(fail(c, ER_E_UNSUPPORTED,
                        "er_create: PME plane %d x %d exceeds one workgroup's LDS (ms2max x ms3max up to about "
                        "128 x 128, or a compiled shape)",
                        prm->ms[1], prm->ms[2]));
    c->numatm = topo->numatm;
    c->nummol = topo->nummol;
    c->ntype = topo->ljtype_max;
    c->numslv = topo->numslv;
    c->prm = *prm;
    c->K1 = prm->ms[0];
    c->K2 = prm->ms[1];
    c->K3 = prm->ms[2];
    c->H = c->K1 / 2 + 1;
    c->ord = prm->splodr;
    c->G = (size_t)c->K1 * c->K2 * c->K3;
    c->Gc = (size_t)c->H * c->K2 * c->K3;
    c->nbx = (c->K1 + ER_BRICK - 1) / ER_BRICK;
    c->nby = (c->K2 + ER_BRICK - 1) / ER_BRICK;
    c->nbz = (c->K3 + brick_z(c->ord) - 1) / brick_z(c->ord);
    c->nbrick = c->nbx * c->nby * c->nbz;
    {
        // bricks ranked along a Morton curve (bit-interleaved bx, by, bz)
        auto spread = [](unsigned v) {
            unsigned r = 0;
            for (int bit = 0; bit < 10; ++bit) r |= ((v >> bit) & 1u) << (3 * bit);
            return r;
        };
        std::vector<std::pair<unsigned, int>> key(c->nbrick);
        for (int z = 0; z < c->nbz; ++z)
            for (int y = 0; y < c->nby; ++y)
                for (int x = 0; x < c->nbx; ++x) {
                    const int id = (z * c->nby + y) * c->nbx + x;
                    key[id] = {spread(x) | (spread(y) << 1) | (spread(z) << 2), id};
                }
        std::sort(key.begin(), key.end());
        std::vector<int> order(c->nbrick);
        for (int k = 0; k < c->nbrick; ++k) order[k] = key[k].second;
        HIPCHK(c, c->d_border.ensure(c->nbrick));
        HIPCHK(c, hipMemcpy(c->d_border.p, order.data(), c->nbrick * sizeof(int), hipMemcpyHostToDevice));
    }
    if (const char *e = getenv("ER_CHUNK")) c->chunk = std::max(NI_B, atoi(e) / NI_B * NI_B);
    if (const char *e = getenv("ER_ECORR_KMIN")) c->ecorr_kmin = std::max(4, std::min(65536, atoi(e) / 4 * 4));

    c->molbeg.resize(c->nummol + 1);
    for (int j = 0; j <= c->nummol; ++j) c->molbeg[j] = topo->mol_begin_index[j] - 1;
    if (c->molbeg[0] != 0 || c->molbeg[c->nummol] != c->numatm)
        return bad(fail(c, ER_E_ARG, "er_create: mol_begin_index inconsistent with numatm"));
    for (int j = 0; j < c->nummol; ++j)
        if (c->molbeg[j + 1] <= c->molbeg[j]) return bad(fail(c, ER_E_ARG, "er_create: empty molecule"));
    std::vector<int> lj(c->numatm);
    for (int a = 0; a < c->numatm; ++a) {
        lj[a] = topo->ljtype[a] - 1;
        if (lj[a] < 0 || lj[a] >= c->ntype) return bad(fail(c, ER_E_ARG, "er_create: ljtype out of range"));
    }
    c->uvspec.assign(topo->uvspec, topo->uvspec + c->nummol);
    c->molq.resize(c->nummol);
    c->qtot = 0.f;
    for (int j = 0; j < c->nummol; ++j) {
        float q = 0.f;  // mol_charge, setconf.F90:838-840
        for (int a = c->molbeg[j]; a < c->molbeg[j + 1]; ++a) q += topo->charge[a];
        c->molq[j] = q;
        c->qtot += q;   // volcorrect total_charge, engproc.F90:1032-1033
    }
    // species ranges: setparam numbers molecules species by species
    c->sp_begin.assign(c->numslv, 0);
    c->sp_end.assign(c->numslv, 0);
    std::vector<int> seen(c->numslv + 1, 0);
    for (int j = 0; j < c->nummol; ++j) {
        const int s = c->uvspec[j];
        if (s < 0 || s > c->numslv) return bad(fail(c, ER_E_ARG, "er_create: uvspec out of range"));
        if (s == 0) continue;
        if (!seen[s]) {
            seen[s] = 1;
            c->sp_begin[s - 1] = j;
        } else if (c->sp_end[s - 1] != j) {
            return bad(fail(c, ER_E_UNSUPPORTED, "er_create: molecules of species %d are not contiguous", s));
        }
        c->sp_end[s - 1] = j + 1;
    }

    // pair parameters
    PairParams &pp = c->pp;
    pp.elecut = prm->elecut;
    pp.lwljcut = prm->lwljcut;
    pp.upljcut = prm->upljcut;
    pp.screen = prm->screen;
    pp.bare = prm->cltype == 0;
    pp.ljswitch = prm->ljswitch;
    pp.lw3 = prm->lwljcut * prm->lwljcut * prm->lwljcut;
    pp.up3 = prm->upljcut * prm->upljcut * prm->upljcut;
    pp.lw6 = pp.lw3 * pp.lw3;
    pp.up6 = pp.up3 * pp.up3;
    gmx_params(12, prm->lwljcut, prm->upljcut, pp.repA, pp.repB, pp.repC);
    gmx_params(6, prm->lwljcut, prm->upljcut, pp.attA, pp.attB, pp.attC);

    HIPCHK(c, hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
    HIPCHK(c, hipStreamCreateWithFlags(&c->stream2, hipStreamNonBlocking));
    HIPCHK(c, hipEventCreateWithFlags(&c->ev_in, hipEventDisableTiming));
    HIPCHK(c, hipEventCreateWithFlags(&c->ev_atoms, hipEventDisableTiming));
    {
        int dev = 0;
        HIPCHK(c, hipGetDevice(&dev));
        HIPCHK(c, hipDeviceGetAttribute(&c->ncu, hipDeviceAttributeMultiprocessorCount, dev));
        // this device's address of the no-allocation scratch, once per device
        if (dev < 0 || dev >= ER_MAX_DEVICES) return bad(fail(c, ER_E_UNSUPPORTED, "er_create: device %d", dev));
        std::lock_guard<std::mutex> lk(g_scratch_mu);
        if (!g_scratch_ptr[dev]) {
            void *p = nullptr;
            const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
            if (e != hipSuccess || !p)
                return bad(fail(c, ER_E_HIP, "er_create: hipGetSymbolAddress(scratch): %s", hipGetErrorString(e)));
            g_scratch_ptr[dev] = static_cast<double *>(p);
        }
        c->scratch_dev = g_scratch_ptr[dev];
    }
    hipStream_t s = c->stream;
    HIPCHK(c, c->d_charge.ensure(c->numatm));
    HIPCHK(c, hipMemcpyAsync(c->d_charge.p, topo->charge, c->numatm * sizeof(float), hipMemcpyHostToDevice, s));
    HIPCHK(c, c->d_ljtype.ensure(c->numatm));
    HIPCHK(c, hipMemcpyAsync(c->d_ljtype.p, lj.data(), c->numatm * sizeof(int), hipMemcpyHostToDevice, s));
    const size_t nt2 = (size_t)c->ntype * c->ntype;
    HIPCHK(c, c->d_ljene.ensure(nt2));
    HIPCHK(c, hipMemcpyAsync(c->d_ljene.p, topo->ljene_mat, nt2 * sizeof(float), hipMemcpyHostToDevice, s));
    HIPCHK(c, c->d_ljlensq.ensure(nt2));
    HIPCHK(c, hipMemcpyAsync(c->d_ljlensq.p, topo->ljlensq_mat, nt2 * sizeof(float), hipMemcpyHostToDevice, s));
    HIPCHK(c, c->d_molbeg.ensure(c->nummol + 1));
    HIPCHK(c, hipMemcpyAsync(c->d_molbeg.p, c->molbeg.data(), (c->nummol + 1) * sizeof(int), hipMemcpyHostToDevice, s));
    HIPCHK(c, c->d_molq.ensure(c->nummol));
    HIPCHK(c, hipMemcpyAsync(c->d_molq.p, c->molq.data(), c->nummol * sizeof(float), hipMemcpyHostToDevice, s));
    HIPCHK(c, c->d_spb.ensure(c->numslv));
    HIPCHK(c, hipMemcpyAsync(c->d_spb.p, c->sp_begin.data(), c->numslv * sizeof(int), hipMemcpyHostToDevice, s));
    HIPCHK(c, c->d_spe.ensure(c->numslv));
    HIPCHK(c, hipMemcpyAsync(c->d_spe.p, c->sp_end.data(), c->numslv * sizeof(int), hipMemcpyHostToDevice, s));

    // PME constants
    for (int d = 0; d < 3; ++d) c->splfc[d] = spline_axis(c->ord, prm->ms[d]);
    HIPCHK(c, c->d_splfc1.ensure(c->K1));
    HIPCHK(c, hipMemcpyAsync(c->d_splfc1.p, c->splfc[0].data(), c->K1 * sizeof(double), hipMemcpyHostToDevice, s));
    HIPCHK(c, c->d_splfc2.ensure(c->K2));
    HIPCHK(c, hipMemcpyAsync(c->d_splfc2.p, c->splfc[1].data(), c->K2 * sizeof(double), hipMemcpyHostToDevice, s));
    HIPCHK(c, c->d_splfc3.ensure(c->K3));
    HIPCHK(c, hipMemcpyAsync(c->d_splfc3.p, c->splfc[2].data(), c->K3 * sizeof(double), hipMemcpyHostToDevice, s));
    HIPCHK(c, upload_tw(c->d_tw1, c->K1, s));
    HIPCHK(c, upload_tw(c->d_tw2, c->K2, s));
    HIPCHK(c, upload_tw(c->d_tw3, c->K3, s));
    HIPCHK(c, upload_tw(c->d_twm, c->K1 / 2, s));
    {
        std::vector<float2> t(ER_TWS_TOTAL);
        for (int r = 4; r <= 16; ++r)
            for (int k = 0; k < r; ++k) {
                const double a = 2.0 * M_PI * (double)k / (double)r;
                t[tws_off(r) + k] = make_float2((float)std::cos(a), (float)std::sin(a));
            }
        HIPCHK(c, c->d_tws.ensure(ER_TWS_TOTAL));
        HIPCHK(c, hipMemcpy(c->d_tws.p, t.data(), ER_TWS_TOTAL * sizeof(float2), hipMemcpyHostToDevice));
    }
    HIPCHK(c, c->d_engfac.ensure(c->Gc));

    // per-atom frame buffers
    HIPCHK(c, c->d_posn.ensure(c->numatm));
    HIPCHK(c, c->d_posu.ensure(c->numatm));
    HIPCHK(c, c->d_splw.ensure((size_t)c->numatm * 3 * c->ord));
    HIPCHK(c, c->d_splg.ensure(c->numatm));
    HIPCHK(c, c->d_molref.ensure(c->nummol));
    HIPCHK(c, c->d_segref.ensure(c->numatm));
    HIPCHK(c, c->d_sitepos_own.ensure((size_t)3 * c->numatm));

    {
        const int rc = slot_state_init(c, c->slots[0]);
        if (rc) return bad(rc);
    }
    // Accumulator layout of the FP64 MFMA: each output register of a known
    // product names its own (row, col); use the MFMA path only if the
    // measured map is a bijection onto the 16x16 tile.
    {
        DBuf<double> probe;
        HIPCHK(c, probe.ensure(256));
        hipLaunchKernelGGL(k_mfma_layout, dim3(1), dim3(64), 0, s, probe.p);
        HIPCHK(c, hipGetLastError());
        std::vector<double> d(256);
        HIPCHK(c, hipMemcpyAsync(d.data(), probe.p, 256 * sizeof(double), hipMemcpyDeviceToHost, s));
        HIPCHK(c, hipStreamSynchronize(s));
        probe.release();
        std::vector<int> map(256), seen(256, 0);
        bool ok = true;
        for (int k = 0; k < 256 && ok; ++k) {
            const double v = d[k];
            const int iv = (int)v;
            if (v != (double)iv || iv < 0 || iv > 255 || seen[iv]) ok = false;
            else {
                seen[iv] = 1;
                map[k] = iv;  // = row + 16*col
            }
        }
        c->use_mfma = ok;
        if (ok) {
            HIPCHK(c, c->d_dmap.ensure(256));
            HIPCHK(c, hipMemcpy(c->d_dmap.p, map.data(), 256 * sizeof(int), hipMemcpyHostToDevice));
        }
    }
    HIPCHK(c, c->d_rmolmax.ensure(1));
    HIPCHK(c, hipStreamSynchronize(s));
    *out = c;
    return ER_OK;
}

extern "C" int er_set_mesh(er_ctx *c, const int32_t *uvmax, const int32_t *uvsoft, const double *uvcrd,
                           int32_t esmax, const double *escrd)
{
    if (!c) return ER_E_ARG;
    ER_NO_CAPTURE(c, "er_set_mesh");
    ++c->alloc_gen;
    if (!uvmax || !uvsoft || !uvcrd) return fail(c, ER_E_ARG, "er_set_mesh: null argument");
    c->uvmax.assign(uvmax, uvmax + c->numslv);
    c->uvsoft.assign(uvsoft, uvsoft + c->numslv);
    c->idmin.assign(c->numslv, 0);
    int tot = 0;
    for (int k = 0; k < c->numslv; ++k) {
        if (c->uvmax[k] <= 0) return fail(c, ER_E_ARG, "er_set_mesh: uvmax must be positive");
        c->idmin[k] = tot;
        tot += c->uvmax[k];
    }
    c->ermax = tot;
    c->uvcrd.assign(uvcrd, uvcrd + tot);
    for (int k = 0; k < c->numslv; ++k)
        for (int b = 1; b < c->uvmax[k]; ++b)
            if (!(c->uvcrd[c->idmin[k] + b - 1] < c->uvcrd[c->idmin[k] + b]))
                return fail(c, ER_E_ARG, "er_set_mesh: mesh of species %d not strictly increasing", k + 1);
    c->esmax = (c->prm.selfcal == 1) ? esmax : 0;
    if (c->prm.selfcal == 1) {
        if (!escrd || esmax <= 0) return fail(c, ER_E_ARG, "er_set_mesh: selfcal needs escrd");
        c->escrd.assign(escrd, escrd + esmax);
    }
    hipStream_t s = c->stream;
    HIPCHK(c, c->d_uvcrd.ensure(tot));
    HIPCHK(c, hipMemcpyAsync(c->d_uvcrd.p, c->uvcrd.data(), tot * sizeof(double), hipMemcpyHostToDevice, s));
    {
        // for FP32 v:  v > c (compared in FP64)  <=>  v > RD32(c)
        std::vector<float> mf(tot);
        for (int b = 0; b < tot; ++b) {
            float f = (float)c->uvcrd[b];
            if ((double)f > c->uvcrd[b]) f = std::nextafter(f, -INFINITY);
            mf[b] = f;
        }
        HIPCHK(c, c->d_meshf.ensure(tot));
        HIPCHK(c, hipMemcpy(c->d_meshf.p, mf.data(), tot * sizeof(float), hipMemcpyHostToDevice));
        // bin brackets: bucket b holds the floats whose order key has top bits b
        std::vector<unsigned short> lut((size_t)c->numslv * ER_LUT_N);
        for (int k = 0; k < c->numslv; ++k) {
            if (c->uvmax[k] > 65535) return fail(c, ER_E_ARG, "er_set_mesh: more than 65535 bins for species %d", k + 1);
            const float *m0 = mf.data() + c->idmin[k], *m1 = m0 + c->uvmax[k];
            for (int b = 0; b < ER_LUT_N - 1; ++b) {
                const uint32_t key = (uint32_t)b << (32 - ER_LUT_BITS);  // smallest order key of the bucket
                const uint32_t bits = (key & 0x80000000u) ? (key & 0x7fffffffu) : ~key;
                float fmin;
                std::memcpy(&fmin, &bits, sizeof fmin);
                // NaN buckets (beyond +inf) take the whole mesh; k_hist never looks them up
                const size_t cnt = (fmin != fmin) ? (size_t)c->uvmax[k] : (size_t)(std::lower_bound(m0, m1, fmin) - m0);
                lut[(size_t)k * ER_LUT_N + b] = (unsigned short)cnt;
            }
            lut[(size_t)k * ER_LUT_N + ER_LUT_N - 1] = (unsigned short)c->uvmax[k];
        }
        HIPCHK(c, c->d_blut.ensure(lut.size()));
        HIPCHK(c, hipMemcpy(c->d_blut.p, lut.data(), lut.size() * sizeof(unsigned short), hipMemcpyHostToDevice));
    }
    HIPCHK(c, c->d_uvmax.ensure(c->numslv));
    HIPCHK(c, hipMemcpyAsync(c->d_uvmax.p, c->uvmax.data(), c->numslv * sizeof(int), hipMemcpyHostToDevice, s));
    HIPCHK(c, c->d_uvsoft.ensure(c->numslv));
    HIPCHK(c, hipMemcpyAsync(c->d_uvsoft.p, c->uvsoft.data(), c->numslv * sizeof(int), hipMemcpyHostToDevice, s));
    HIPCHK(c, c->d_idmin.ensure(c->numslv));
    HIPCHK(c, hipMemcpyAsync(c->d_idmin.p, c->idmin.data(), c->numslv * sizeof(int), hipMemcpyHostToDevice, s));
    HIPCHK(c, c->d_escrd.ensure(std::max(1, c->esmax)));
    if (c->esmax > 0)
        HIPCHK(c, hipMemcpyAsync(c->d_escrd.p, c->escrd.data(), c->esmax * sizeof(double), hipMemcpyHostToDevice, s));
    c->ermax = tot;
    for (Slot &sl : c->slots) {
        const int rc = slot_acc_alloc(c, sl);
        if (rc) return rc;
    }
    HIPCHK(c, c->d_epart.ensure((size_t)ER_KSLICES * tot));
    if (c->prm.corrcal == 1 && c->use_mfma) HIPCHK(c, c->d_part.ensure((size_t)ER_KSLICES * tot * tot));
    HIPCHK(c, hipStreamSynchronize(s));
    c->have_mesh = true;
    return er_clear(c);
}

static int set_slot_solute(er_ctx *c, Slot &sl, int32_t nsite, const int32_t *ljtype, const float *charge)
{
    if (nsite <= 0 || !ljtype || !charge) return fail(c, ER_E_ARG, "er_set_solute: bad arguments");
    std::vector<int> lj(nsite);
    float q = 0.f;
    for (int s = 0; s < nsite; ++s) {
        lj[s] = ljtype[s] - 1;
        if (lj[s] < 0 || lj[s] >= c->ntype) return fail(c, ER_E_ARG, "er_set_solute: ljtype out of range");
        q += charge[s];
    }
    sl.qslt = q;
    if (sl.nslt != nsite) sl.have_ins = false;  // insertions of another shape are void
    sl.nslt = nsite;
    HIPCHK(c, sl.d_sltq.ensure(nsite));
    HIPCHK(c, hipMemcpy(sl.d_sltq.p, charge, nsite * sizeof(float), hipMemcpyHostToDevice));
    HIPCHK(c, sl.d_sltlj.ensure(nsite));
    HIPCHK(c, hipMemcpy(sl.d_sltlj.p, lj.data(), nsite * sizeof(int), hipMemcpyHostToDevice));
    return ER_OK;
}

extern "C" int er_set_solute(er_ctx *c, int32_t nsite, const int32_t *ljtype, const float *charge)
{
    if (!c) return ER_E_ARG;
    ER_NO_CAPTURE(c, "er_set_solute");
    ++c->alloc_gen;
    if (c->prm.slttype == 1) return fail(c, ER_E_STATE, "er_set_solute: soln runs use er_set_solute_molecules");
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return set_slot_solute(c, c->S(), nsite, ljtype, charge);
}

extern "C" int er_set_solutes(er_ctx *c, int32_t nsolute, const int32_t *nsite, const int32_t *ljtype,
                              const float *charge)
{
    if (!c) return ER_E_ARG;
    ER_NO_CAPTURE(c, "er_set_solutes");
    ++c->alloc_gen;
    if (c->prm.slttype == 1) return fail(c, ER_E_STATE, "er_set_solutes: only for refs runs");
    if (nsolute <= 0 || !nsite || !ljtype || !charge) return fail(c, ER_E_ARG, "er_set_solutes: bad arguments");
    HIPCHK(c, hipStreamSynchronize(c->stream));
    for (size_t k = (size_t)nsolute; k < c->slots.size(); ++k) c->slots[k].release();
    c->slots.resize(nsolute);
    c->cur = 0;
    size_t off = 0;
    for (int k = 0; k < nsolute; ++k) {
        Slot &sl = c->slots[k];
        int rc = set_slot_solute(c, sl, nsite[k], ljtype + off, charge + off);
        if (!rc) rc = slot_state_init(c, sl);
        if (!rc && c->have_mesh) rc = slot_acc_alloc(c, sl);
        if (!rc && c->have_mesh) rc = slot_clear(c, sl);
        if (rc) return rc;
        sl.have_ins = false;
        off += (size_t)nsite[k];
    }
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return ER_OK;
}

extern "C" int er_select_solute(er_ctx *c, int32_t k)
{
    if (!c) return ER_E_ARG;
    if (k < 1 || k > (int)c->slots.size()) return fail(c, ER_E_ARG, "er_select_solute: %d out of 1..%d", k,
                                                       (int)c->slots.size());
    c->cur = k - 1;
    return ER_OK;
}

extern "C" int er_set_solute_molecules(er_ctx *c, int32_t nslt, const int32_t *tagslt)
{
    if (!c) return ER_E_ARG;
    ER_NO_CAPTURE(c, "er_set_solute_molecules");
    ++c->alloc_gen;
    if (c->prm.slttype != 1) return fail(c, ER_E_STATE, "er_set_solute_molecules: only for soln runs");
    if (nslt <= 0 || !tagslt) return fail(c, ER_E_ARG, "er_set_solute_molecules: bad arguments");
    c->tagslt.resize(nslt);
    int ns = -1;
    for (int k = 0; k < nslt; ++k) {
        const int m = tagslt[k] - 1;
        if (m < 0 || m >= c->nummol) return fail(c, ER_E_ARG, "er_set_solute_molecules: index out of range");
        const int sz = c->molbeg[m + 1] - c->molbeg[m];
        if (ns >= 0 && sz != ns) return fail(c, ER_E_ARG, "set_slt: solute molecules differ in size");
        ns = sz;
        c->tagslt[k] = m;
    }
    c->nsltmol = nslt;
    c->S().nslt = ns;
    HIPCHK(c, hipStreamSynchronize(c->stream));  // a run in flight may still read the list
    HIPCHK(c, c->d_tag.ensure(nslt));
    HIPCHK(c, hipMemcpy(c->d_tag.p, c->tagslt.data(), nslt * sizeof(int), hipMemcpyHostToDevice));
    return ER_OK;
}

// Morton keys of nc cells per axis: coordinates below the next power of two
// p interleave into keys below p^3 (32^3 at most)
static int morton_keys(int nc)
{
    int p = 1;
    while (p < nc) p <<= 1;
    return p * p * p;
}

// realcal_prepare's cell handling (realcal.F90:776-896) + get_inverted_cell
static int frame_cell(er_ctx *c, const float *cell)
{
    if (c->prm.boxshp == 0) {
        // open system (realcal_prepare without normalize_periodic): no cell.
        // Zero cell constants leave wrap_site an identity and min_image a
        // sign change that keeps every norm; splines and the cell list see
        // fractional coordinates 0 (they serve no kernel of this path).
        FrameConst &fo = c->fc;
        std::memset(&fo, 0, sizeof fo);
        fo.is_cuboid = 1;
        c->volume_d = 0.0;
        for (int k = 0; k < 9; ++k) c->invcl_d[k] = 0.0;
        c->nc = 1;
        c->cell_w = 1.0e30f;
        c->cell_changed = false;
        std::memcpy(c->cell, cell, sizeof c->cell);
        return ER_OK;
    }
    // column-major: cell(i,j) = cell[i + 3j]; sub-diagonal (2,1),(3,1),(3,2).
    // cn is cell_normal in FP32 as the reference's real; cd follows it in
    // FP64 (the same rotation, flip and lattice shifts) for the lattice
    // coordinates of the pair energies.
    float cn[9];
    double cd[9], qd[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1};
    std::memcpy(cn, cell, sizeof cn);
    for (int k = 0; k < 9; ++k) cd[k] = cell[k];
    FrameConst &fc = c->fc;
    fc.rotate = 0;
    if (cell[1] != 0.f || cell[2] != 0.f || cell[5] != 0.f) {
        // make_cell_uppertriangular (realcal.F90:776-817): cell = QR by
        // modified Gram-Schmidt, cell_normal = R, pair positions Q^T x.  Each
        // column is normalised by its own length (the reference divides all
        // three by the length of column 1, line 799, which is no rotation).
        double t[9], q[9], r[9] = {0};
        for (int k = 0; k < 9; ++k) t[k] = cell[k];
        for (int i = 0; i < 3; ++i) {
            const double len = std::sqrt(t[3 * i] * t[3 * i] + t[3 * i + 1] * t[3 * i + 1] + t[3 * i + 2] * t[3 * i + 2]);
            if (!(len > 0.0)) return fail(c, ER_E_ARG, "cell has a zero or dependent axis");
            for (int k = 0; k < 3; ++k) q[k + 3 * i] = t[k + 3 * i] / len;
            r[i + 3 * i] = len;
            for (int j = i + 1; j < 3; ++j) {
                const double d = q[3 * i] * t[3 * j] + q[3 * i + 1] * t[3 * j + 1] + q[3 * i + 2] * t[3 * j + 2];
                r[i + 3 * j] = d;
                for (int k = 0; k < 3; ++k) t[k + 3 * j] -= d * q[k + 3 * i];
            }
        }
        for (int k = 0; k < 9; ++k) cn[k] = (float)r[k];
        for (int k = 0; k < 9; ++k) cd[k] = r[k];
        for (int i = 0; i < 3; ++i)
            for (int k = 0; k < 3; ++k) {
                fc.rotT[3 * i + k] = (float)q[k + 3 * i];
                qd[3 * i + k] = q[k + 3 * i];
            }
        fc.rotate = 1;
    }
    // normalize_cell_vector (realcal.F90:826-830): only column 1 can be flipped
    for (int i = 0; i < 3; ++i)
        if (cn[0] < 0.f)
            for (int r = 0; r < 3; ++r) {
                cn[r + 3 * i] = -cn[r + 3 * i];
                cd[r + 3 * i] = -cd[r + 3 * i];
            }
    // its lattice reduction (realcal.F90:832-847): a column whose (1,2),
    // (1,3) or (2,3) element exceeds half the diagonal by more than 1e-4 is
    // shifted by the nearest whole multiple of column 1 (or 2), in float as
    // the reference's real.  The reference scales by invcell_normal as the
    // previous frame left it (normalize_periodic sets it afterwards); this
    // uses 1/cell_normal(i,i) of the frame itself, which is the same value
    // for every frame after the first of a constant cell.
    bool reduced = false;
    {
        auto anint = [](float x) { return x >= 0.f ? std::floor(x + 0.5f) : -std::floor(-x + 0.5f); };
        const float thres = 1.0e-4f;
        const float inv0 = 1.0f / std::fabs(cn[0]);
        if (std::fabs(cn[3]) > 0.5f * cn[0] + thres) {
            const float f = anint(cn[3] * inv0);
            for (int r = 0; r < 3; ++r) cn[r + 3] -= cn[r] * f;
            for (int r = 0; r < 3; ++r) cd[r + 3] -= cd[r] * f;
            reduced = true;
        }
        if (std::fabs(cn[6]) > 0.5f * cn[0] + thres) {
            const float f = anint(cn[6] * inv0);
            for (int r = 0; r < 3; ++r) cn[r + 6] -= cn[r] * f;
            for (int r = 0; r < 3; ++r) cd[r + 6] -= cd[r] * f;
            reduced = true;
        }
        const float inv1 = 1.0f / std::fabs(cn[4]);
        if (std::fabs(cn[7]) > 0.5f * cn[4] + thres) {
            const float f = anint(cn[7] * inv1);
            for (int r = 0; r < 3; ++r) cn[r + 6] -= cn[r + 3] * f;
            for (int r = 0; r < 3; ++r) cd[r + 6] -= cd[r + 3] * f;
            reduced = true;
        }
    }
    const float cutoff = std::min(c->prm.elecut, c->prm.upljcut);
    for (int i = 0; i < 3; ++i)
        if (cutoff > cn[i + 3 * i] * 0.5f)
            return fail(c, ER_E_ARG,
                        "One of axis in periodic cell is too small. This is either the box is too small "
                        "compared to the cell, or the periodic cell is too skewed.");
    std::memcpy(fc.cellN, cn, sizeof cn);
    for (int i = 0; i < 3; ++i) {
        fc.lenN[i] = std::fabs(cn[i + 3 * i]);
        fc.invN[i] = 1.0f / fc.lenN[i];
    }
    // lattice coordinates: s = cd^-1 Q^T x (cd upper triangular), and back
    // to Cartesian differences through cell_normal / 2^32
    {
        const double a = cd[0], b = cd[3], d = cd[6], e = cd[4], f = cd[7], g = cd[8];
        const double ic[9] = {1.0 / a, 0, 0, -b / (a * e), 1.0 / e, 0, (b * f - d * e) / (a * e * g), -f / (e * g),
                              1.0 / g};
        for (int k = 0; k < 3; ++k)
            for (int j = 0; j < 3; ++j) {
                double m = 0.0;
                for (int r = 0; r < 3; ++r) m += ic[k + 3 * r] * qd[3 * r + j];
                fc.fracM[k + 3 * j] = m;
            }
        for (int k = 0; k < 9; ++k) fc.dsc[k] = (float)std::ldexp(cd[k], -32);
    }
    fc.is_cuboid = !(std::fabs(cell[3]) > 1e-8f || std::fabs(cell[6]) > 1e-8f || std::fabs(cell[7]) > 1e-8f);
    // inverse cell in double (engproc.F90:1049-1070)
    double m[9];
    for (int k = 0; k < 9; ++k) m[k] = cell[k];
#define M(i, j) m[(i) + 3 * (j)]
    const double v = M(0, 0) * M(1, 1) * M(2, 2) + M(0, 1) * M(1, 2) * M(2, 0) + M(0, 2) * M(1, 0) * M(2, 1) -
                     M(0, 2) * M(1, 1) * M(2, 0) - M(0, 1) * M(1, 0) * M(2, 2) - M(0, 0) * M(1, 2) * M(2, 1);
    double *iv = c->invcl_d;
    iv[0 + 3 * 0] = M(1, 1) * M(2, 2) - M(1, 2) * M(2, 1);
    iv[0 + 3 * 1] = M(0, 2) * M(2, 1) - M(0, 1) * M(2, 2);
    iv[0 + 3 * 2] = M(0, 1) * M(1, 2) - M(0, 2) * M(1, 1);
    iv[1 + 3 * 0] = M(1, 2) * M(2, 0) - M(1, 0) * M(2, 2);
    iv[1 + 3 * 1] = M(0, 0) * M(2, 2) - M(0, 2) * M(2, 0);
    iv[1 + 3 * 2] = M(0, 2) * M(1, 0) - M(0, 0) * M(1, 2);
    iv[2 + 3 * 0] = M(1, 0) * M(2, 1) - M(1, 1) * M(2, 0);
    iv[2 + 3 * 1] = M(0, 1) * M(2, 0) - M(0, 0) * M(2, 1);
    iv[2 + 3 * 2] = M(0, 0) * M(1, 1) - M(0, 1) * M(1, 0);
#undef M
    for (int k = 0; k < 9; ++k) {
        iv[k] /= v;
        fc.invcl[k] = (float)iv[k];
        fc.invcld[k] = iv[k];
        fc.invP[k] = fc.invcl[k];
    }
    if (fc.rotate || reduced) {  // inverse of the upper-triangular cell_normal the pairs use
        const double a = cn[0], b = cn[3], d = cn[6], e = cn[4], f = cn[7], g = cn[8];
        double ip[9] = {1.0 / a, 0, 0, -b / (a * e), 1.0 / e, 0, (b * f - d * e) / (a * e * g), -f / (e * g), 1.0 / g};
        for (int k = 0; k < 9; ++k) fc.invP[k] = (float)ip[k];
    }
    c->volume_d = v;
    fc.volume = (float)v;
    // cell-list resolution: nc cells per axis of >= 3.5 A perpendicular
    // width, in the cell the pair positions are wrapped into (cell_normal)
    {
        double wmin = 1e300, pm[9];
        for (int k = 0; k < 9; ++k) pm[k] = cn[k];
        const double pv = std::fabs(pm[0] * pm[4] * pm[8]);  // upper triangular
        for (int k = 0; k < 3; ++k) {
            const int i1 = (k + 1) % 3, i2 = (k + 2) % 3;
            const double *a = pm + 3 * i1, *b = pm + 3 * i2;
            const double cx = a[1] * b[2] - a[2] * b[1], cy = a[2] * b[0] - a[0] * b[2], cz = a[0] * b[1] - a[1] * b[0];
            wmin = std::min(wmin, pv / std::sqrt(cx * cx + cy * cy + cz * cz));
        }
        c->cell_wmin = wmin;
        c->nc = std::max(1, std::min(32, (int)(wmin / (c->cell_min > 0.f ? c->cell_min : 3.5f))));
        c->ncell = morton_keys(c->nc);
        c->cell_w = (float)(wmin / c->nc);
    }
    c->cell_changed = !c->have_prev_cell || std::memcmp(c->prev_cell, cell, sizeof c->prev_cell) != 0;
    std::memcpy(c->cell, cell, sizeof c->cell);
    return ER_OK;
}

extern "C" int er_set_frame(er_ctx *c, const float *sitepos, const float *cell, float stat_weight_system)
{
    if (!c) return ER_E_ARG;
    ER_NO_CAPTURE(c, "er_set_frame (use er_set_frame_device)");
    if (!sitepos || !cell) return fail(c, ER_E_ARG, "er_set_frame: null argument");
    int rc = frame_cell(c, cell);
    if (rc) return rc;
    // the previous run may still read the buffer
    HIPCHK(c, hipStreamSynchronize(c->stream));
    HIPCHK(c, hipMemcpy(c->d_sitepos_own.p, sitepos, (size_t)3 * c->numatm * sizeof(float), hipMemcpyHostToDevice));
    c->d_sitepos = c->d_sitepos_own.p;
    c->w_sys = stat_weight_system;
    c->have_frame = true;
    c->frame_dirty = true;
    return ER_OK;
}

extern "C" int er_set_frame_device(er_ctx *c, const float *sitepos_dev, const float *cell, float stat_weight_system)
{
    if (!c) return ER_E_ARG;
    if (!sitepos_dev || !cell) return fail(c, ER_E_ARG, "er_set_frame_device: null argument");
    int rc = frame_cell(c, cell);
    if (rc) return rc;
    c->d_sitepos = sitepos_dev;
    c->w_sys = stat_weight_system;
    c->have_frame = true;
    c->frame_dirty = true;
    return ER_OK;
}

static int ins_common(er_ctx *c, int32_t nins, const float *wslt, bool copies)
{
    if (c->prm.slttype == 1) return fail(c, ER_E_STATE, "insertions are for refs runs");
    if (c->S().nslt <= 0) return fail(c, ER_E_STATE, "er_set_solute must precede insertions");
    if (nins <= 0) return fail(c, ER_E_ARG, "nins must be positive");
    if (c->capturing && (copies || wslt))
        return fail(c, ER_E_STATE, "a recorded graph takes er_set_insertions_device without stat_weight_solute");
    // a run still in flight may read the buffers about to be overwritten
    if (copies || wslt) HIPCHK(c, hipStreamSynchronize(c->stream));
    if (c->capturing && c->S().d_wslt.n < (size_t)nins)
        return fail(c, ER_E_STATE, "er_graph_begin: run these insertions once before recording");
    c->S().nins = nins;
    {
        const float *before = c->S().d_wslt.p;
        HIPCHK(c, c->S().d_wslt.ensure(nins));
        if (c->S().d_wslt.p != before) ++c->alloc_gen;
    }
    c->S().wslt_given = wslt != nullptr;
    if (wslt) HIPCHK(c, hipMemcpy(c->S().d_wslt.p, wslt, nins * sizeof(float), hipMemcpyHostToDevice));
    c->S().have_ins = true;
    return ER_OK;
}

extern "C" int er_set_insertions(er_ctx *c, int32_t nins, const float *sltpos, const float *stat_weight_solute)
{
    if (!c) return ER_E_ARG;
    if (!sltpos) return fail(c, ER_E_ARG, "er_set_insertions: null sltpos");
    int rc = ins_common(c, nins, stat_weight_solute, true);
    if (rc) return rc;
    const size_t n = (size_t)3 * c->S().nslt * nins;
    HIPCHK(c, c->S().d_sltpos_own.ensure(n));  // read only by the runs' copy of the coordinates
    HIPCHK(c, hipMemcpy(c->S().d_sltpos_own.p, sltpos, n * sizeof(float), hipMemcpyHostToDevice));
    c->S().d_sltpos = c->S().d_sltpos_own.p;
    return ER_OK;
}

extern "C" int er_set_insertions_device(er_ctx *c, int32_t nins, const float *sltpos_dev,
                                        const float *stat_weight_solute)
{
    if (!c) return ER_E_ARG;
    if (!sltpos_dev) return fail(c, ER_E_ARG, "er_set_insertions_device: null sltpos");
    int rc = ins_common(c, nins, stat_weight_solute, false);
    if (rc) return rc;
    c->S().d_sltpos = sltpos_dev;
    return ER_OK;
}

// ---- kernel dispatch on compile-time grid sizes and spline order ----

// line split N = R1 * R2 of the supported lengths for the four-step plane FFT
template <int K> struct Split;
template <> struct Split<32> { static constexpr int A = 8, B = 4; };
template <> struct Split<36> { static constexpr int A = 9, B = 4; };
template <> struct Split<40> { static constexpr int A = 8, B = 5; };
template <> struct Split<48> { static constexpr int A = 8, B = 6; };
template <> struct Split<64> { static constexpr int A = 8, B = 8; };
template <> struct Split<72> { static constexpr int A = 8, B = 9; };
template <> struct Split<80> { static constexpr int A = 8, B = 10; };
template <> struct Split<96> { static constexpr int A = 8, B = 12; };
template <> struct Split<108> { static constexpr int A = 12, B = 9; };
template <> struct Split<128> { static constexpr int A = 8, B = 16; };

// the plane kernel over insertions [n_off, n_off + nch) of the run, whose T
// rows start at insertion t_off; sparse: the scatter + forward DFT form (the
// solutes of more than ER_SPARSE_SITES sites)
template <int K2, int K3> static hipError_t launch_spec2d_k(er_ctx *c, int n_off, int nch, int t_off, bool sparse)
{
    constexpr int R1 = Split<K3>::A, R2 = Split<K3>::B, C1 = Split<K2>::A, C2 = Split<K2>::B;
    constexpr int NT = (K2 * K3 >= 5184) ? 512 : 256;  // large planes: more threads, fewer values per thread
    const size_t lds = (size_t)K2 * (K3 + 1) * sizeof(float2) + (size_t)ER_SCH * (K2 + K3) * sizeof(float2) +
                       (K2 + K3) * sizeof(float2) + (NT / 64) * sizeof(double);
    static bool attr = false;
    if (!attr && lds > 65536) {
        (void)hipFuncSetAttribute((const void *)k_pme_spec2d<K2, K3, R1, R2, C1, C2, NT, false>,
                                  hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
        (void)hipFuncSetAttribute((const void *)k_pme_spec2d<K2, K3, R1, R2, C1, C2, NT, true>,
                                  hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
        attr = true;
    }
    const int nslot = (int)c->plan.desc.size();
    float2 *T = c->d_T.p + (size_t)(n_off - t_off) * c->H * K2 * K3;  // the launch's first T row
    // many sites: scatter + forward plane DFT beats nslt * K2 * K3 products
    if (sparse)
        hipLaunchKernelGGL((k_pme_spec2d<K2, K3, R1, R2, C1, C2, NT, true>), dim3(c->H, nch), dim3(NT), lds,
                           c->stream, nslot, c->d_sd.p, c->ord, c->K1, n_off, c->d_Ftab.p, c->d_insw.p,
                           c->d_insg.p, c->d_engfac.p, c->d_tw2.p, c->d_tw3.p, c->d_tws.p, T, c->d_selfpart.p);
    else
        hipLaunchKernelGGL((k_pme_spec2d<K2, K3, R1, R2, C1, C2, NT, false>), dim3(c->H, nch), dim3(NT), lds,
                           c->stream, nslot, c->d_sd.p, c->ord, c->K1, n_off, c->d_Ftab.p, c->d_insw.p,
                           c->d_insg.p, c->d_engfac.p, c->d_tw2.p, c->d_tw3.p, c->d_tws.p, T, c->d_selfpart.p);
    return hipGetLastError();
}

// the direct-DFT plane kernel for shapes without a compiled one
static hipError_t launch_spec2d_gen(er_ctx *c, int n_off, int nch, int t_off, bool sparse)
{
    static bool attr = false;
    if (!attr) {
        (void)hipFuncSetAttribute((const void *)k_pme_spec2d_gen, hipFuncAttributeMaxDynamicSharedMemorySize,
                                  160 * 1024);
        attr = true;
    }
    const size_t lds = spec2d_gen_lds(c->K2, c->K3);
    float2 *T = c->d_T.p + (size_t)(n_off - t_off) * c->Gc;
    hipLaunchKernelGGL(k_pme_spec2d_gen, dim3(c->H, nch), dim3(ER_GEN_NT), lds, c->stream,
                       (int)c->plan.desc.size(), c->d_sd.p, c->ord, c->K1, c->K2, c->K3, (int)sparse, n_off,
                       c->d_Ftab.p, c->d_insw.p, c->d_insg.p, c->d_engfac.p, c->d_tw2.p, c->d_tw3.p, T,
                       c->d_selfpart.p);
    return hipGetLastError();
}

static hipError_t launch_spec2d(er_ctx *c, int n_off, int nch, int t_off, bool sparse)
{
    if (nch <= 0) return hipSuccess;
    if (!plane_ok(c->K2, c->K3)) return launch_spec2d_gen(c, n_off, nch, t_off, sparse);
    if (c->K2 == c->K3) {
        switch (c->K2) {
        case 32: return launch_spec2d_k<32, 32>(c, n_off, nch, t_off, sparse);
        case 36: return launch_spec2d_k<36, 36>(c, n_off, nch, t_off, sparse);
        case 40: return launch_spec2d_k<40, 40>(c, n_off, nch, t_off, sparse);
        case 48: return launch_spec2d_k<48, 48>(c, n_off, nch, t_off, sparse);
        case 64: return launch_spec2d_k<64, 64>(c, n_off, nch, t_off, sparse);
        case 72: return launch_spec2d_k<72, 72>(c, n_off, nch, t_off, sparse);
        case 80: return launch_spec2d_k<80, 80>(c, n_off, nch, t_off, sparse);
        case 96: return launch_spec2d_k<96, 96>(c, n_off, nch, t_off, sparse);
        case 108: return launch_spec2d_k<108, 108>(c, n_off, nch, t_off, sparse);
        case 128: return launch_spec2d_k<128, 128>(c, n_off, nch, t_off, sparse);
        }
    } else if (c->K2 == 128 && c->K3 == 108) {
        return launch_spec2d_k<128, 108>(c, n_off, nch, t_off, sparse);
    }
    return hipErrorInvalidValue;
}

static int divisor_le(int K, int cap)
{
    for (int r = std::min(K, cap); r >= 1; --r)
        if (K % r == 0) return r;
    return 1;
}

template <int K1, int NI> static hipError_t launch_axis1_k(er_ctx *c, int nch)
{
    if constexpr (NI == 16) {
        // transposed-store form: 8 (or 4) lines per workgroup
        const int ntile = (nch + NI - 1) / NI;
        auto go = [&](auto rl) -> hipError_t {
            constexpr int RL = decltype(rl)::value;
            const size_t lds = (size_t)RL * (K1 * 16 + 8) * sizeof(float);
            static bool attr = false;
            if (!attr && lds > 65536) {
                (void)hipFuncSetAttribute((const void *)k_pme_axis1_t<K1, RL>,
                                          hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
                attr = true;
            }
            const size_t nblk = (size_t)(c->K3 / RL) * c->K2 * ntile;  // one flat, XCD-remapped range
            hipLaunchKernelGGL((k_pme_axis1_t<K1, RL>), dim3((unsigned)(8 * ((nblk + 7) / 8))), dim3(RL * 16), lds,
                               c->stream, nch, c->K2, c->K3, c->d_T.p, c->d_phi.p, c->d_tw1.p, c->d_twm.p);
            return hipGetLastError();
        };
        if (c->K3 % 8 == 0) return go(std::integral_constant<int, 8>{});
        if (c->K3 % 4 == 0) return go(std::integral_constant<int, 4>{});
    }
    const int R = divisor_le(c->K3, 256 / NI);  // x3 lines per block
    const int nt = ((R * NI + 63) / 64) * 64;
    const int ntile = (nch + NI - 1) / NI;
    hipLaunchKernelGGL((k_pme_axis1<K1, NI>), dim3(c->K3 / R, c->K2, ntile), dim3(nt), 0, c->stream, nch, c->K2,
                       c->K3, R, c->d_T.p, c->d_phi.p, c->d_tw1.p, c->d_twm.p);
    return hipGetLastError();
}

template <int NI> static hipError_t launch_axis1(er_ctx *c, int nch)
{
    if (!fft_len_ok(c->K1)) {  // the direct C2R kernel
        const int ntile = (nch + NI - 1) / NI;
        const size_t lds = ((size_t)NI * c->H + c->K1) * sizeof(float2);
        hipLaunchKernelGGL(k_pme_axis1_gen, dim3(c->K3, c->K2, ntile), dim3(ER_GEN_NT), lds, c->stream, nch, NI,
                           c->K1, c->K2, c->K3, c->d_T.p, c->d_phi.p, c->d_tw1.p);
        return hipGetLastError();
    }
    switch (c->K1) {
    case 32: return launch_axis1_k<32, NI>(c, nch);
    case 36: return launch_axis1_k<36, NI>(c, nch);
    case 40: return launch_axis1_k<40, NI>(c, nch);
    case 48: return launch_axis1_k<48, NI>(c, nch);
    case 64: return launch_axis1_k<64, NI>(c, nch);
    case 72: return launch_axis1_k<72, NI>(c, nch);
    case 80: return launch_axis1_k<80, NI>(c, nch);
    case 96: return launch_axis1_k<96, NI>(c, nch);
    case 108: return launch_axis1_k<108, NI>(c, nch);
    case 128: return launch_axis1_k<128, NI>(c, nch);
    }
    return hipErrorInvalidValue;
}

template <int ORD, int NI, int LPM> static hipError_t launch_gather_k(er_ctx *c, int n_off, int nch)
{
    constexpr int MPW = 64 / LPM;
    constexpr int NXCD = 8;  // MI355X: 8 XCDs, workgroups dispatched round-robin
    const int mols_per_block = 4 * MPW;
    int ngx = (c->nummol + mols_per_block - 1) / mols_per_block;
    ngx = (ngx + NXCD - 1) / NXCD * NXCD;
    const int ngy = (nch + LPM - 1) / LPM;
    const bool soln = c->prm.slttype == 1;
    hipLaunchKernelGGL((k_gather<ORD, NI, LPM>), dim3(ngx * ngy), dim3(256), 0, c->stream, c->nummol, nch, n_off,
                       c->K1, c->K2, c->K3, c->d_perm.p, ngx, NXCD, c->d_molbeg.p, c->d_posn.p, c->d_splw.p,
                       c->d_splg.p, c->d_molq.p, c->d_phi.p, soln ? c->d_tag.p : (const int *)nullptr,
                       (int)c->plan.desc.size(), c->d_sd.p, c->prm.screen, c->fc.volume, c->d_Ucat.p);
    return hipGetLastError();
}

template <int NI, int LPM> static hipError_t launch_gather(er_ctx *c, int n_off, int nch)
{
    switch (c->ord) {
    case 3: return launch_gather_k<3, NI, LPM>(c, n_off, nch);
    case 4: return launch_gather_k<4, NI, LPM>(c, n_off, nch);
    case 5: return launch_gather_k<5, NI, LPM>(c, n_off, nch);
    case 6: return launch_gather_k<6, NI, LPM>(c, n_off, nch);
    case 7: return launch_gather_k<7, NI, LPM>(c, n_off, nch);
    case 8: return launch_gather_k<8, NI, LPM>(c, n_off, nch);
    case 9: return launch_gather_k<9, NI, LPM>(c, n_off, nch);
    case 10: return launch_gather_k<10, NI, LPM>(c, n_off, nch);
    }
    return hipErrorInvalidValue;
}

// spline orders with a brick gather: a region row of RS = 8 + ORD - 1 grid
// points x 16 insertions must be one wave instruction (RS <= 16) and the
// region must fit the 160 KB LDS (order 8: 15 x 15 x 11 points, 158 KB)
static bool brick_order(int ord) { return ord >= 3 && ord <= 8; }

template <int ORD, int NI> static hipError_t launch_gather_brick_k(er_ctx *c, int n_off, int nch)
{
    static_assert((size_t)(ER_BRICK + ORD - 1) * (ER_BRICK + ORD - 1) * (brick_z(ORD) + ORD - 1) * (NI / 4) * 16 <=
                      160 * 1024,
                  "the staged region fits the LDS");
    constexpr int NXCD = 8;
    constexpr int RS = ER_BRICK + ORD - 1, RZ = brick_z(ORD) + ORD - 1;
    const size_t lds = (size_t)RS * RS * RZ * (NI / 4) * 16;  // the staged region
    static bool attr = false;
    if (!attr) {
        (void)hipFuncSetAttribute((const void *)k_gather_brick<ORD, NI>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                  (int)lds);
        attr = true;
    }
    const int ntile = (nch + NI - 1) / NI;
    const size_t nwg = (size_t)NXCD * ((c->nbrick + NXCD - 1) / NXCD) * ntile;
    hipLaunchKernelGGL((k_gather_brick<ORD, NI>), dim3((unsigned)nwg), dim3(512), lds, c->stream, c->K1, c->K2, c->K3,
                       c->nbx, c->nby, c->nbrick, c->d_bdesc.p, nch, c->d_brec.p, c->d_phi.p, c->numatm,
                       c->d_Eat.p);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    const bool soln = c->prm.slttype == 1;
    const int mpb = NI == 16 ? ER_MR_MPT4 * 64 : ER_MR_MPT * (256 / NI);  // molecules per block
    hipLaunchKernelGGL((k_mol_reduce<NI>), dim3((unsigned)((c->nummol + mpb - 1) / mpb), ntile), dim3(256),
                       0, c->stream, c->nummol,
                       nch, n_off, c->numatm, c->d_molbeg.p, c->d_molq.p, c->d_Eat.p,
                       soln ? c->d_tag.p : (const int *)nullptr, (int)c->plan.desc.size(), c->d_sd.p, c->prm.screen,
                       c->fc.volume, c->d_Ucat.p);
    return hipGetLastError();
}

// order 4: the streaming gather, one workgroup per CU (its two LDS buffers
// fill a CU), a whole number of workgroups per XCD
template <int NI> static hipError_t launch_gather_stream_k(er_ctx *c, int n_off, int nch)
{
    const int nwg = 8 * std::max(1, c->ncu / 8);
    hipLaunchKernelGGL((k_gather_stream<4, NI>), dim3((unsigned)nwg), dim3(ER_GS_NT), 0, c->stream, c->K1, c->K2,
                       c->K3, c->nbx, c->nby, c->nbrick, c->d_bdesc.p, nch, c->d_brec.p, c->d_phi.p, c->numatm,
                       c->d_Eat.p);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    const bool soln = c->prm.slttype == 1;
    const int ntile = (nch + NI - 1) / NI;
    const int mpb = NI == 16 ? ER_MR_MPT4 * 64 : ER_MR_MPT * (256 / NI);  // molecules per block
    hipLaunchKernelGGL((k_mol_reduce<NI>), dim3((unsigned)((c->nummol + mpb - 1) / mpb), ntile), dim3(256), 0,
                       c->stream, c->nummol, nch, n_off, c->numatm, c->d_molbeg.p, c->d_molq.p, c->d_Eat.p,
                       soln ? c->d_tag.p : (const int *)nullptr, (int)c->plan.desc.size(), c->d_sd.p, c->prm.screen,
                       c->fc.volume, c->d_Ucat.p);
    return hipGetLastError();
}

template <int NI> static hipError_t launch_hist(er_ctx *c, int n_off, int nch)
{
    // 512 threads per tile and slice for the batched path (more waves per CU
    // than 256 at the same slice count and global merge traffic); the plain
    // path's single-insertion tiles keep 256 (its per-thread partials are
    // reduced by one thread)
    const int NT = NI > 1 ? 512 : 256;
    const size_t lds = (size_t)NT * sizeof(double) + (size_t)c->ermax * sizeof(float) +
                       (size_t)NI * (c->ermax | 1) * sizeof(uint32_t) + (size_t)NT * 2 * sizeof(float) +
                       (size_t)c->numslv * ER_LUT_N * sizeof(unsigned short);
    if (lds > 160 * 1024) return hipErrorInvalidValue;
    if (lds > 65536)
        (void)hipFuncSetAttribute((const void *)k_hist<NI>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    const bool soln = c->prm.slttype == 1;
    HistArgs ha{c->nummol, nch, n_off, c->ermax, c->numslv, c->prm.slttype, c->nslice};
    const int ntile = (nch + NI - 1) / NI;
    hipLaunchKernelGGL((k_hist<NI>), dim3(ntile, c->nslice), dim3(NT), lds, c->stream, ha,
                       (int)c->plan.desc.size(), c->d_sd.p, c->d_Ucat.p,
                       soln ? c->d_tag.p : (const int *)nullptr, c->d_meshf.p, c->d_blut.p, c->d_spb.p, c->d_spe.p,
                       c->d_idmin.p, c->d_uvmax.p, c->d_uvsoft.p, c->d_C.p,
                       soln ? c->d_flpart.p : (double *)nullptr);
    return hipGetLastError();
}

// PPPM coefficients: gf_b of calc_gfb_pppm (recpcal.F90:179-199; the last
// update of gf_b(1) sees the inner loop's exit value l = 0) and the alias
// range of recpcal.F90:228-233, formed in single precision as there
static PPPMCoef pppm_coef(const er_ctx *c)
{
    PPPMCoef pc{};
    const int P = c->ord;
    pc.order = P;
    double b[ER_MAXORD] = {1.0};
    for (int m = 1; m <= P - 1; ++m) {
        for (int l = m; l >= 1; --l)
            b[l] = 4.0 * (b[l] * (l - m) * (l - m - 0.5) - b[l - 1] * (l - m - 1) * (l - m - 1));
        b[0] = 4.0 * (b[0] * (0 - m) * (0 - m - 0.5));
    }
    double fact = 1.0;
    for (int n = 2; n <= 2 * P - 1; ++n) fact *= n;
    for (int m = 0; m < P; ++m) pc.b[m] = b[m] / fact;
    const float root = std::pow(-std::log(1.0e-7f), 0.25f), pi = 3.14159265358979f;
    const float mm[3] = {c->prm.screen / (pi * c->K1) * root, c->prm.screen / (pi * c->K2) * root,
                         c->prm.screen / (pi * c->K3) * root};
    int a[3];
    for (int d = 0; d < 3; ++d)
        a[d] = (int)(c->cell[0 + 3 * d] * mm[0] + c->cell[1 + 3 * d] * mm[1] + c->cell[2 + 3 * d] * mm[2]);
    pc.a1 = std::max(a[0], 0);
    pc.a2 = std::max(a[1], 0);
    pc.a3 = std::max(a[2], 0);
    return pc;
}

#define HIPCHK_E(x)                              \
    do {                                         \
        const hipError_t e_ = (x);               \
        if (e_ != hipSuccess) return e_;         \
    } while (0)

// the main stream waits for the frame's atom preparation (once per frame)
static hipError_t join_atoms(er_ctx *c)
{
    if (!c->atoms_pending) return hipSuccess;
    c->atoms_pending = false;
    return hipStreamWaitEvent(c->stream, c->ev_atoms, 0);
}

static hipError_t prep_frame(er_ctx *c)
{
    const int B = 256;
    if (c->capturing && c->capture_runs == 0 && c->cell_changed) c->capture_engfac = true;
    if (c->cell_changed && c->prm.cltype == 0) {  // bare Coulomb: no influence function
        std::memcpy(c->prev_cell, c->cell, sizeof c->cell);
        c->have_prev_cell = true;
        c->cell_changed = false;
    }
    if (c->cell_changed && c->prm.cltype == 3) {
        const double *i = c->invcl_d;
        hipLaunchKernelGGL(k_engfac_pppm, dim3((unsigned)((c->Gc + B - 1) / B)), dim3(B), 0, c->stream, c->K1,
                           c->K2, c->K3, pppm_coef(c), i[0], i[1], i[2], i[3], i[4], i[5], i[6], i[7], i[8],
                           c->volume_d, (double)c->prm.screen, c->d_engfac.p);
        std::memcpy(c->prev_cell, c->cell, sizeof c->cell);
        c->have_prev_cell = true;
        c->cell_changed = false;
    }
    if (c->cell_changed) {
        const double *i = c->invcl_d;
        hipLaunchKernelGGL(k_engfac, dim3((unsigned)((c->Gc + B - 1) / B)), dim3(B), 0, c->stream, c->K1, c->K2, c->K3,
                           c->ord, c->d_splfc1.p, c->d_splfc2.p, c->d_splfc3.p, i[0], i[1], i[2], i[3], i[4], i[5],
                           i[6], i[7], i[8], c->volume_d, (double)c->prm.screen, c->d_engfac.p);
        std::memcpy(c->prev_cell, c->cell, sizeof c->cell);
        c->have_prev_cell = true;
        c->cell_changed = false;
    }
    if (c->frame_dirty) {
        // after the frame's upload on the main stream; while a graph is
        // recorded the work stays on the main stream
        hipStream_t s2 = c->capturing ? c->stream : c->stream2;
        if (!c->capturing) {
            HIPCHK_E(hipEventRecord(c->ev_in, c->stream));
            HIPCHK_E(hipStreamWaitEvent(s2, c->ev_in, 0));
        }
        // five launches: atoms (and the counters cleared), molecules, both
        // counting passes, both scans, then the brick records with the
        // molecules' scatter into cell order
        hipLaunchKernelGGL(k_prep_atoms, dim3((c->numatm + B - 1) / B), dim3(B), 0, s2, c->numatm, c->d_sitepos,
                           c->d_charge.p, c->fc, c->K1, c->K2, c->K3, c->ord, c->d_posn.p, c->d_posu.p, c->d_splw.p,
                           c->d_splg.p, c->d_ccnt.p, c->ncell, c->d_bcnt.p, c->nbrick, c->d_rmolmax.p);
        hipLaunchKernelGGL(k_prep_mols, dim3((c->nummol + B - 1) / B), dim3(B), 0, s2, c->nummol,
                           c->d_molbeg.p, c->d_posn.p, c->fc, c->d_molref.p, c->d_rmolmax.p, c->d_segref.p);
        // spatial order of the molecules (the cell list) and of the atoms by PME brick
        const int nsort = c->nummol + c->numatm;
        hipLaunchKernelGGL(k_sort_count, dim3((nsort + B - 1) / B), dim3(B), 0, s2, c->nummol, c->d_molref.p,
                           c->fc, c->nc, c->d_ccnt.p, c->d_keyrank.p, c->numatm, c->d_splg.p, c->K1, c->K2, c->K3,
                           c->nbx, c->nby, brick_z(c->ord), c->d_bcnt.p, c->d_akeyrank.p);
        hipLaunchKernelGGL(k_sort_scan, dim3(2), dim3(1024), 0, s2, c->ncell, c->d_ccnt.p, c->d_cstart.p, c->nbrick,
                           c->d_bcnt.p, c->d_bstart.p);
        switch (c->ord) {
#define ER_BREC(O)                                                                                              \
    case O:                                                                                                     \
        hipLaunchKernelGGL(k_brick_records<O>, dim3((nsort + B - 1) / B), dim3(B), 0, s2, c->numatm,            \
                           c->d_akeyrank.p, c->d_bstart.p, c->d_posn.p, c->d_splw.p, c->d_splg.p, c->K1, c->K2, \
                           c->K3, c->nbx, c->nby, c->d_brec.p, c->nummol, c->d_keyrank.p, c->d_cstart.p,        \
                           c->d_perm.p, c->nbrick, c->d_border.p, c->d_bdesc.p);                                \
        break;
            ER_BREC(3) ER_BREC(4) ER_BREC(5) ER_BREC(6) ER_BREC(7) ER_BREC(8)
#undef ER_BREC
        default:  // orders 9 and 10: the per-molecule gather, no records
            hipLaunchKernelGGL(k_cell_scatter, dim3((c->nummol + B - 1) / B), dim3(B), 0, s2, c->nummol,
                               c->d_keyrank.p, c->d_cstart.p, c->d_perm.p);
            break;
        }
        if (!c->capturing) {
            HIPCHK_E(hipEventRecord(c->ev_atoms, s2));
            c->atoms_pending = true;
        }
        c->frame_dirty = false;
    }
    return hipGetLastError();
}

// a run buffer of at least `count` elements: grown as needed, except in an
// er_run_noalloc call or while a graph is recorded, where a buffer too small
// is an error (nothing is allocated)
template <class T> static int want(er_ctx *c, DBuf<T> &b, size_t count, const char *name)
{
    if (!c->noalloc && !c->capturing) {
        const T *before = b.p;
        HIPCHK(c, b.ensure(count));
        if (b.p != before) ++c->alloc_gen;
        return ER_OK;
    }
    if (b.p && count <= b.n) return ER_OK;
    return fail(c, ER_E_STATE,
                "%s: buffer %s holds %zu elements, this run needs %zu; an er_run of the same solutes and at least "
                "as many insertions (same launch shapes) must size it first",
                c->capturing ? "er_graph_begin" : "er_run_noalloc", name, b.n, count);
}

// insertion slices of the split-K edens/ecorr sums over a solute's nch
// insertions: slices of kper (a multiple of 4, at least ecorr_kmin)
// insertions.  Every slice writes a partial tile set of ermax^2 that the
// reduce reads back, while one slice adds straight into ecorr: a solute
// whose 64 x 64 tiles alone give every CU a block takes one slice, a
// smaller one as many as give its (tile, slice) blocks about four per CU (at
// most ER_KSLICES).  The count depends on the solute alone, so a solute's
// sums are the same whatever else shares its run.
static void ecorr_slices(const er_ctx *c, int nch, int &kper, int &nks)
{
    const long nb = (c->ermax + 63) / 64;
    const long tiles = std::max(1L, nb * (nb + 1) / 2);
    const long ncu = std::max(1, c->ncu);
    const int want =
        tiles >= ncu ? 1 : (int)std::max(1L, std::min<long>(ER_KSLICES, (4 * ncu + tiles - 1) / tiles));
    kper = std::max(c->ecorr_kmin, (((nch + want - 1) / want) + 3) / 4 * 4);
    nks = (nch + kper - 1) / kper;
}

// workgroups per insertion in the pair search: the slot's tuned value in a
// batched run (the shape it was tuned for), else enough to fill the chip (at
// least 4 per CU); a many-site solute reaches far, so its candidates are
// spread over 16.  Every (insertion, molecule) sum has one owner whatever
// the split, so the split never changes a result.
static int pair_split(const Slot &sl, int nins, bool batched)
{
    if (batched && sl.pair_split > 0) return sl.pair_split;
    if (sl.nslt > 8) return 16;
    return std::max(1, std::min(16, (4 * 256 + nins - 1) / nins));
}

// The run's solutes end to end: the solutes of few sites first (dense plane
// spectra), then the many-site ones (sparse spectra), each solute's
// insertion range padded to the interleave NI.
static int build_plan(er_ctx *c, const std::vector<int> &todo, int mode)
{
    auto &P = c->plan;
    const bool soln = c->prm.slttype == 1;
    const bool batched = mode == ER_BATCHED;
    P.NI = batched ? NI_B : 1;
    P.order.clear();
    for (int pass = 0; pass < 2; ++pass)
        for (int k : todo)
            if ((c->slots[k].nslt > ER_SPARSE_SITES) == (pass == 1)) P.order.push_back(k);
    P.desc.assign(P.order.size(), SlotDesc{});
    long long ins = 0, site = 0, blk = 0;
    int z = 0, wrow = 1;
    P.ndense = 0;
    for (size_t i = 0; i < P.order.size(); ++i) {
        Slot &sl = c->slots[P.order[i]];
        const int nins = soln ? c->nsltmol : sl.nins;
        SlotDesc &d = P.desc[i];
        d.nslt = sl.nslt;
        d.ins_off = (int)ins;
        d.nins = nins;
        d.site_off = (int)site;
        d.pair_split = pair_split(sl, batched ? nins : 1, batched);
        d.pair_blk = (int)blk;
        ecorr_slices(c, nins, d.kper, d.nks);
        d.slice_off = z;
        d.qslt = sl.qslt;
        d.sltq = sl.d_sltq.p;
        d.sltlj = sl.d_sltlj.p;
        d.wslt = sl.wslt_given ? sl.d_wslt.p : nullptr;
        d.seq = sl.d_seq.p;
        d.state = sl.d_state.p;
        d.flags = sl.d_flags.p;
        d.fstate = sl.d_fstate.p;
        d.minmax = sl.d_minmax.p;
        d.edens = sl.d_edens.p;
        d.ecorr = sl.d_ecorr.p;
        d.eself = sl.d_eself.p;
        d.slnuv = sl.d_slnuv.p;
        d.errkey = sl.d_errkey.p;
        d.warnkey = sl.d_warnkey.p;
        d.halt = sl.d_halt.p;
        ins += (nins + P.NI - 1) / P.NI * P.NI;
        site += (long long)nins * sl.nslt;
        blk += (long long)nins * d.pair_split;
        z += d.nks;
        wrow = std::max(wrow, (nins + 255) / 256);
        if (sl.nslt <= ER_SPARSE_SITES) P.ndense = (int)ins;
    }
    if (ins >= (1ll << 30) || site >= (1ll << 30) || blk >= (1ll << 31) - 1)
        return fail(c, ER_E_UNSUPPORTED, "er_run: %lld insertions, %lld sites in one run exceed this build's range",
                    ins, site);
    P.ntot = (int)ins;
    P.nsite = (int)site;
    P.pair_blocks = (int)blk;
    P.nz = z;
    P.wrow = wrow;
    return ER_OK;
}

// the descriptors on the device; an upload waits for the runs in flight,
// which read the previous ones (the descriptors change only with the shapes)
static int upload_plan(er_ctx *c)
{
    const auto &d = c->plan.desc;
    if (d.size() == c->sd_up.size() && std::memcmp(d.data(), c->sd_up.data(), d.size() * sizeof(SlotDesc)) == 0)
        return ER_OK;
    if (c->capturing)
        return fail(c, ER_E_STATE,
                    "er_graph_begin: the solutes, insertion counts or launch shapes differ from the last er_run; "
                    "run them once before recording");
    int rc = want(c, c->d_sd, d.size(), "descriptors");
    if (rc) return rc;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    HIPCHK(c, hipMemcpy(c->d_sd.p, d.data(), d.size() * sizeof(SlotDesc), hipMemcpyHostToDevice));
    c->sd_up = d;
    return ER_OK;
}

// the split-K partials (ecorr, edens) and the weights' block partials of
// the run: the context's buffers, or in an er_run_noalloc call consecutive
// ranges of the scratch array
struct AccPartials {
    double *part, *epart, *wpart;
};

static void partial_sizes(const er_ctx *c, bool batched, size_t &npart, size_t &nepart, size_t &nwpart)
{
    const auto &P = c->plan;
    int kper1, nks1;
    ecorr_slices(c, 1, kper1, nks1);
    const size_t nz = batched ? (size_t)P.nz : (size_t)nks1;
    const bool mfma = c->prm.corrcal == 1 && c->use_mfma;
    npart = mfma ? nz * c->ermax * c->ermax : 0;
    nepart = nz * c->ermax;
    nwpart = batched ? (size_t)4 * P.desc.size() * P.wrow : 4;
}

static AccPartials acc_partials(er_ctx *c, bool batched)
{
    if (!c->noalloc) return {c->d_part.p, c->d_epart.p, c->d_wpart.p};
    size_t np, ne, nw;
    partial_sizes(c, batched, np, ne, nw);
    double *p = c->scratch_dev;
    return {p, p + np, p + np + ne};
}

static int ensure_run_buffers(er_ctx *c, int chunk)
{
    const auto &P = c->plan;
    const int NI = P.NI;
    const bool soln = c->prm.slttype == 1;
    const size_t nsite = (size_t)P.nsite, ntot = (size_t)P.ntot;
    int rc;
#define ER_WANT(buf, count)                                      \
    do {                                                         \
        if ((rc = want(c, buf, (size_t)(count), #buf)) != ER_OK) \
            return rc;                                           \
    } while (0)
    // the partial sums first: a run beyond the no-allocation scratch is refused outright
    {
        size_t np, ne, nw;
        partial_sizes(c, NI > 1, np, ne, nw);
        if (c->noalloc) {
            if (np + ne + nw > ER_NOALLOC_SCRATCH_BYTES / sizeof(double))
                return fail(c, ER_E_UNSUPPORTED,
                            "er_run_noalloc: the run's partial sums (%d insertions, %d bins) need %zu bytes, above "
                            "the %zu-byte scratch (use er_run)",
                            P.ntot, c->ermax, (np + ne + nw) * sizeof(double), (size_t)ER_NOALLOC_SCRATCH_BYTES);
        } else {
            // at least the slices set_mesh sized, so small runs never shrink them
            ER_WANT(c->d_wpart, nw);
            ER_WANT(c->d_epart, std::max(ne, (size_t)ER_KSLICES * c->ermax));
            if (np) ER_WANT(c->d_part, std::max(np, (size_t)ER_KSLICES * c->ermax * c->ermax));
        }
    }
    ER_WANT(c->d_insposn, nsite);
    ER_WANT(c->d_insposu, nsite);
    ER_WANT(c->d_insw, nsite * 3 * c->ord);
    ER_WANT(c->d_insg, nsite);
    if (!soln) ER_WANT(c->d_sltcat, nsite * 3);
    ER_WANT(c->d_Ftab, nsite * (c->H + c->K2 + c->K3));
    ER_WANT(c->d_insref, ntot);
    ER_WANT(c->d_selfreal, ntot);
    ER_WANT(c->d_resself, ntot);
    ER_WANT(c->d_selfpart, ntot * c->H);
    ER_WANT(c->d_T, (size_t)chunk * c->Gc);
    const size_t ntile = (size_t)(chunk + NI - 1) / NI;
    ER_WANT(c->d_phi, ntile * NI * c->G);
    ER_WANT(c->d_Ucat, ntot * c->nummol);
    ER_WANT(c->d_u0cat, ntot);
    ER_WANT(c->d_C, ntot * c->ermax);
    ER_WANT(c->d_w, ntot);
    ER_WANT(c->d_perm, c->nummol);
    ER_WANT(c->d_keyrank, 2 * (size_t)c->nummol);
    ER_WANT(c->d_ccnt, c->ncell);
    ER_WANT(c->d_cstart, c->ncell + 1);
    if (brick_order(c->ord)) ER_WANT(c->d_brec, (size_t)c->numatm * brick_rec_floats(c->ord));
    ER_WANT(c->d_akeyrank, 2 * (size_t)c->numatm);
    ER_WANT(c->d_bcnt, c->nbrick);
    ER_WANT(c->d_bstart, c->nbrick + 1);
    ER_WANT(c->d_bdesc, c->nbrick);
    if (NI > 1 && brick_order(c->ord)) ER_WANT(c->d_Eat, ntile * NI * (size_t)c->numatm);
    // histogram slices: soln runs keep one built-in rule (their flceng partial
    // sums follow it) from the chunk's tiles; refs runs take about 2048
    // (tile, slice) workgroups over the tiles of one histogram launch (the
    // whole run when batched, one insertion when plain), but no slice with
    // fewer than NI ermax / 4 molecules (every workgroup zeroes and flushes
    // NI x ermax bins) unless the CUs would otherwise go short of
    // workgroups; a tuned count wins
    const int ntile_chunk = (chunk + NI - 1) / NI;
    if (c->prm.slttype == 1) {
        c->nslice = std::max(1, std::min(64, (2048 + ntile_chunk - 1) / ntile_chunk));
    } else {
        const long ntile_launch = NI > 1 ? std::max(1L, (long)((ntot + NI - 1) / NI)) : 1L;
        const long by_fill = (2048 + ntile_launch - 1) / ntile_launch;
        const long by_bins = std::max(1L, 4L * c->nummol / ((long)NI * std::max(1, c->ermax)));
        const long by_cus = (std::max(1, c->ncu) + ntile_launch - 1) / ntile_launch;
        c->nslice = (int)std::max(1L, std::min(64L, std::min(by_fill, std::max(by_bins, by_cus))));
    }
    if (c->hist_slices > 0 && c->prm.slttype != 1) c->nslice = c->hist_slices;
    if (soln) {
        ER_WANT(c->d_flceng, ntot * c->numslv);
        ER_WANT(c->d_flpart, ntot * c->nslice * c->numslv);
    }
#undef ER_WANT
    return ER_OK;
}

// reciprocal part of insertions [n_off, n_off+nch) of the run (a chunk; T
// rows start at n_off): plane spectra (few-site solutes, then many-site
// ones), axis-1 transform, gather and molecule sums into U; NI = 16 for the
// batched path, 1 for the plain path
template <int NI> static hipError_t run_chunk(er_ctx *c, int n_off, int nch)
{
    hipError_t e;
    {
        ProfScope ps(c, ER_K_SPEC2D);
        const int nd = std::max(0, std::min(n_off + nch, c->plan.ndense) - n_off);  // few-site insertions
        if ((e = launch_spec2d(c, n_off, nd, n_off, false)) != hipSuccess) return e;
        if ((e = launch_spec2d(c, n_off + nd, nch - nd, n_off, true)) != hipSuccess) return e;
    }
    {
        ProfScope ps(c, ER_K_AXIS1);
        if ((e = launch_axis1<NI>(c, nch)) != hipSuccess) return e;
    }
    if ((e = join_atoms(c)) != hipSuccess) return e;
    ProfScope ps(c, ER_K_GATHER);
    if constexpr (NI == 1) return launch_gather<1, 1>(c, n_off, nch);
    else {
        switch (c->ord) {
        case 3: return launch_gather_brick_k<3, NI>(c, n_off, nch);
        case 4: return launch_gather_stream_k<NI>(c, n_off, nch);
        case 5: return launch_gather_brick_k<5, NI>(c, n_off, nch);
        case 6: return launch_gather_brick_k<6, NI>(c, n_off, nch);
        case 7: return launch_gather_brick_k<7, NI>(c, n_off, nch);
        case 8: return launch_gather_brick_k<8, NI>(c, n_off, nch);
        default: return launch_gather<NI, NI>(c, n_off, nch);  // 9, 10: regions beyond one wave's row
        }
    }
}

// the real-space pairs (workgroups [blk0, blk0 + nblk) of the run; open
// systems: insertions [n_off, n_off + nch)) and the histograms of
// insertions [n_off, n_off + nch)
template <int NI> static hipError_t run_pairs_hist(er_ctx *c, int n_off, int nch, int blk0, int nblk)
{
    hipError_t e;
    if ((e = join_atoms(c)) != hipSuccess) return e;
    const bool soln = c->prm.slttype == 1;
    const int nslot = (int)c->plan.desc.size();
    PairArgs pa{};
    pa.nmol = c->nummol;
    pa.nchunk = nch;
    pa.n_off = n_off;
    pa.NI = NI;
    pa.ntype = c->ntype;
    pa.nc = c->nc;
    pa.pp = c->pp;
    pa.cell_w = c->cell_w;
    {
        ProfScope ps(c, ER_K_PAIRS);
        if (c->prm.boxshp == 0) {  // open system: every molecule, no cell list; y = insertions of a chunk
            const int step = std::max(NI, std::min(32768, c->chunk) / NI * NI);
            for (int s0 = 0; s0 < nch; s0 += step) {
                pa.n_off = n_off + s0;
                pa.nchunk = std::min(step, nch - s0);
                hipLaunchKernelGGL(k_pairs_open, dim3((unsigned)((c->nummol + 255) / 256), (unsigned)pa.nchunk),
                                   dim3(256), 0, c->stream, pa, nslot, c->d_sd.p, c->d_molbeg.p, c->d_posn.p,
                                   c->d_ljtype.p, c->d_insposn.p, soln ? c->d_tag.p : (const int *)nullptr,
                                   c->d_ljene.p, c->d_ljlensq.p, c->d_Ucat.p);
                if ((e = hipGetLastError()) != hipSuccess) return e;
            }
        } else if (nblk > 0) {
            // the few-site and many-site kernels, for the classes of the
            // solutes whose blocks [pair_blk, pair_blk + nins * pair_split)
            // meet [blk0, blk0 + nblk); each returns from the other's blocks
            bool few = false, many = false;
            for (const SlotDesc &d : c->plan.desc) {
                const long b0 = d.pair_blk, b1 = (long)d.pair_blk + (long)d.nins * d.pair_split;
                if (b1 > blk0 && b0 < (long)blk0 + nblk) (d.nslt > 8 ? many : few) = true;
            }
            for (int k = 0; k < 2; ++k) {
                if (!(k == 0 ? few : many)) continue;
                hipLaunchKernelGGL(k == 0 ? k_pairs : k_pairs_many, dim3((unsigned)nblk), dim3(256), 0, c->stream,
                                   pa, blk0, nslot, c->d_sd.p, c->fc, c->d_molbeg.p, c->d_posu.p, c->d_ljtype.p,
                                   c->d_molref.p, c->d_segref.p, c->d_rmolmax.p, c->d_cstart.p, c->d_perm.p,
                                   c->d_insposn.p, c->d_insposu.p, c->d_insref.p,
                                   soln ? c->d_tag.p : (const int *)nullptr, c->d_ljene.p, c->d_ljlensq.p,
                                   c->d_Ucat.p);
                if ((e = hipGetLastError()) != hipSuccess) return e;
            }
        }
    }
    ProfScope ps(c, ER_K_HIST);
    return launch_hist<NI>(c, n_off, nch);
}

// weights, edens and ecorr (after the histograms); update_histogram's
// accumulation part over the insertions `ar` names (every solute of the run,
// or one insertion range of one solute)
static hipError_t run_accumulate(er_ctx *c, AccRange ar)
{
    const bool soln = c->prm.slttype == 1;
    const bool all = ar.nch < 0;
    const auto &P = c->plan;
    const int nslot = (int)P.desc.size();
    const int ns = all ? nslot : 1;  // solutes of this accumulation (grid dimension)
    const AccPartials ap = acc_partials(c, all);
    const SlotDesc *sd = c->d_sd.p;
    {
        ProfScope ps(c, ER_K_WEIGHTS);
        WeightArgs wa{};
        wa.H = (c->prm.cltype == 0) ? 0 : c->H;   // no reciprocal self part for bare Coulomb
        wa.bare = c->prm.cltype == 0;
        wa.slttype = c->prm.slttype;
        wa.wgtslf = c->prm.wgtslf;
        wa.estype = c->prm.estype;
        wa.selfcal = c->prm.selfcal;
        wa.esmax = c->esmax;
        wa.rigid = c->prm.slttype == 2;
        wa.numslv = c->numslv;
        wa.nslice = c->nslice;
        wa.w_sys = c->w_sys;
        wa.volume = c->fc.volume;
        wa.screen = c->prm.screen;
        wa.qtot = c->qtot;
        wa.temp = c->prm.temp;
        const int wrow = all ? P.wrow : std::max(1, (ar.nch + 255) / 256);
        hipLaunchKernelGGL(k_weights_init, dim3(ns), dim3(64), 0, c->stream, wa, ar, sd, c->d_selfpart.p,
                           c->d_selfreal.p, c->d_resself.p);
        hipLaunchKernelGGL(k_weights, dim3(wrow, ns), dim3(256), 0, c->stream, wa, ar, sd, c->d_selfpart.p,
                           c->d_selfreal.p, c->d_resself.p, c->d_u0cat.p, c->d_w.p,
                           soln ? c->d_flpart.p : (const double *)nullptr, soln ? c->d_flceng.p : (double *)nullptr,
                           ap.wpart);
        hipLaunchKernelGGL(k_weights_fin, dim3(ns), dim3(256), 0, c->stream, wa, ar, sd, wrow, ap.wpart,
                           c->d_u0cat.p, c->d_w.p, c->d_escrd.p, soln ? c->d_flpart.p : (const double *)nullptr,
                           soln ? c->d_flceng.p : (const double *)nullptr);
    }
    const int nz = all ? P.nz : (ar.nch + ar.kper - 1) / ar.kper;
    const bool mfma = c->prm.corrcal == 1 && c->use_mfma;
    {
        ProfScope ps(c, ER_K_ECORR);
        const int nb = (c->ermax + 63) / 64;
        // k_ecorr: (slice, upper-triangle tile) pairs, a whole number per XCD
        const long ntri = (long)nb * (nb + 1) / 2;
        const long nwg = 8 * ((ntri * nz + 7) / 8);
        if (mfma)
            hipLaunchKernelGGL(k_ecorr, dim3((unsigned)nwg), dim3(256), 0, c->stream, ar, nslot, sd, c->ermax, nz,
                               c->d_C.p, c->d_w.p, c->d_dmap.p, ap.part, ap.epart);
        else
            hipLaunchKernelGGL(k_edens, dim3((c->ermax + 255) / 256, nz), dim3(256), 0, c->stream, ar, nslot, sd,
                               c->ermax, c->d_C.p, c->d_w.p, ap.epart);
        // k_ecorr adds a one-slice solute straight into ecorr and edens: when
        // every solute of the run has one slice, the reduce has nothing to add
        bool reduce = !mfma;
        if (mfma && all)
            for (const auto &d : P.desc) reduce = reduce || d.nks > 1;
        else if (mfma)
            reduce = nz > 1;
        if (reduce)
            hipLaunchKernelGGL(k_ecorr_reduce, dim3((c->ermax + 255) / 256, mfma ? c->ermax : 1, ns), dim3(256), 0,
                               c->stream, ar, sd, c->ermax, (int)mfma, ap.part, ap.epart);
        if (c->prm.corrcal == 1 && !c->use_mfma) {
            const int nb16 = (c->ermax + 15) / 16;
            hipLaunchKernelGGL(k_ecorr_valu, dim3(nb16, nb16, ns), dim3(16, 16), 0, c->stream, ar, sd, c->ermax,
                               c->d_C.p, c->d_w.p);
        }
    }
    return hipGetLastError();
}

// one run over the solutes `todo` of the frame: per-insertion preparation,
// the reciprocal part in chunks, pairs, histograms and accumulation, each a
// single launch over every solute (the plain path: one insertion at a time)
static int run_slots(er_ctx *c, const std::vector<int> &todo, int mode)
{
    const bool soln = c->prm.slttype == 1;
    const bool batched = mode == ER_BATCHED;
    for (int k : todo) {
        const int nins = soln ? c->nsltmol : c->slots[k].nins;
        if (nins >= (1 << 21) || c->nummol >= (1 << 21) - 1)
            return fail(c, ER_E_UNSUPPORTED, "er_run: at most 2^21-1 insertions per solute and 2^21-2 molecules");
    }
    int rc = build_plan(c, todo, mode);
    if (rc) return rc;
    const auto &P = c->plan;
    const int NI = P.NI;
    const int nslot = (int)P.desc.size();
    const int chunk = batched ? c->chunk : 1;
    if ((rc = ensure_run_buffers(c, chunk))) return rc;
    if ((rc = upload_plan(c))) return rc;
    const SlotDesc *sd = c->d_sd.p;
    {
        ProfScope ps(c, ER_K_PREP);
        HIPCHK(c, prep_frame(c));
        if (!soln)  // the insertions' coordinates, solutes end to end
            for (int i = 0; i < nslot; ++i) {
                const Slot &sl = c->slots[P.order[i]];
                const size_t n = (size_t)P.desc[i].nins * sl.nslt * 3;
                HIPCHK(c, hipMemcpyAsync(c->d_sltcat.p + (size_t)P.desc[i].site_off * 3, sl.d_sltpos,
                                         n * sizeof(float), hipMemcpyDeviceToDevice, c->stream));
            }
        hipLaunchKernelGGL(k_prep_sites, dim3((unsigned)((std::max(P.nsite, nslot) + 255) / 256)), dim3(256), 0,
                           c->stream, P.nsite,
                           nslot, sd, soln ? (const float *)nullptr : c->d_sltcat.p, c->d_sitepos, c->d_charge.p,
                           c->d_molbeg.p, soln ? c->d_tag.p : (const int *)nullptr, c->fc, c->K1, c->K2, c->K3,
                           c->ord, c->d_insposn.p, c->d_insposu.p, c->d_insw.p, c->d_insg.p);
        HIPCHK(c, hipGetLastError());
        hipLaunchKernelGGL(k_prep_ins, dim3((unsigned)std::max(1, P.ntot)), dim3(64 * ER_PI_WAVES), 0, c->stream, P.ntot, nslot,
                           sd, c->fc, c->prm.screen, (int)(c->prm.cltype == 0), c->d_insposn.p, c->d_insref.p,
                           c->d_selfreal.p, c->d_resself.p);
        HIPCHK(c, hipGetLastError());
        const size_t nf = (size_t)P.nsite * (c->H + c->K2 + c->K3);
        hipLaunchKernelGGL(k_ins_factors, dim3((unsigned)((nf + 255) / 256)), dim3(256), 0, c->stream, P.nsite,
                           c->ord, c->K1, c->K2, c->K3, c->H, c->d_insposn.p, c->d_insw.p, c->d_insg.p, c->d_tw1.p,
                           c->d_tw2.p, c->d_tw3.p, c->d_Ftab.p);
        HIPCHK(c, hipGetLastError());
        const size_t nC = (size_t)P.ntot * c->ermax;
        hipLaunchKernelGGL(k_zero_u32, dim3((unsigned)((nC + 255) / 256)), dim3(256), 0, c->stream, nC, c->d_C.p);
        HIPCHK(c, hipGetLastError());
        if (c->prm.cltype == 0) {
            // bare Coulomb (engproc.F90:761-769, 819-826): no reciprocal
            // part; U starts at zero for the pair sums
            const size_t n = (size_t)P.ntot * c->nummol;
            hipLaunchKernelGGL(k_zero_u32, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, c->stream, n,
                               reinterpret_cast<uint32_t *>(c->d_Ucat.p));
            HIPCHK(c, hipGetLastError());
        }
    }
    if (batched) {
        if (c->prm.cltype != 0)
            for (int n0 = 0; n0 < P.ntot; n0 += chunk) HIPCHK(c, run_chunk<NI_B>(c, n0, std::min(chunk, P.ntot - n0)));
        HIPCHK(c, run_pairs_hist<NI_B>(c, 0, P.ntot, 0, P.pair_blocks));
        HIPCHK(c, run_accumulate(c, AccRange{0, 0, -1, 0}));
    } else {
        // one insertion per launch sequence; ER_PLAIN also waits for each,
        // like the reference's per-insertion host loop
        int kper1, nks1;
        ecorr_slices(c, 1, kper1, nks1);
        for (int i = 0; i < nslot; ++i) {
            const SlotDesc &d = P.desc[i];
            for (int j = 0; j < d.nins; ++j) {
                const int n = d.ins_off + j;
                if (c->prm.cltype != 0) HIPCHK(c, run_chunk<1>(c, n, 1));
                HIPCHK(c, run_pairs_hist<1>(c, n, 1, d.pair_blk + j * d.pair_split, d.pair_split));
                HIPCHK(c, run_accumulate(c, AccRange{i, j, 1, kper1}));
                if (mode == ER_PLAIN) HIPCHK(c, hipStreamSynchronize(c->stream));
            }
        }
    }
    HIPCHK(c, join_atoms(c));  // a run without chunks still orders the next upload after the atom work
    if (c->capturing) ++c->capture_runs;
    hipLaunchKernelGGL(k_halt_capture, dim3(nslot), dim3(64), 0, c->stream, sd, NI, c->nummol, c->d_Ucat.p,
                       c->d_u0cat.p);
    HIPCHK(c, hipGetLastError());
    for (int i = 0; i < nslot; ++i) {
        Slot &sl = c->slots[P.order[i]];
        sl.U_view = c->d_Ucat.p + (size_t)P.desc[i].ins_off * c->nummol;
        sl.u0_view = c->d_u0cat.p + P.desc[i].ins_off;
        sl.last_NI = NI;
        sl.last_nins = P.desc[i].nins;
    }
    // solutes outside this run keep no energies (their rows were reused)
    for (size_t k = 0; k < c->slots.size(); ++k)
        if (std::find(todo.begin(), todo.end(), (int)k) == todo.end()) c->slots[k].last_nins = 0;
    return ER_OK;
}

// the slots a run processes (every solute with insertions; the one soln
// slot), after the state checks of er_run
static int run_plan(er_ctx *c, const char *who, std::vector<int> &todo)
{
    if (!c->have_mesh) return fail(c, ER_E_STATE, "%s: er_set_mesh first", who);
    if (!c->have_frame) return fail(c, ER_E_STATE, "%s: no frame set", who);
    todo.clear();
    if (c->prm.slttype == 1) {
        if (c->nsltmol <= 0) return fail(c, ER_E_STATE, "%s: er_set_solute_molecules first", who);
        todo.push_back(0);
    } else {
        for (int k = 0; k < (int)c->slots.size(); ++k)
            if (c->slots[k].have_ins) todo.push_back(k);
        if (todo.empty()) return fail(c, ER_E_STATE, "%s: no insertions set", who);
    }
    return ER_OK;
}

extern "C" int er_run(er_ctx *c, int32_t mode)
{
    if (!c) return ER_E_ARG;
    if (mode != ER_PLAIN && mode != ER_BATCHED && mode != ER_PLAIN_ASYNC) return fail(c, ER_E_ARG, "er_run: mode");
    if (c->capturing && mode != ER_BATCHED)
        return fail(c, ER_E_STATE, "er_run: a graph records ER_BATCHED runs only (ER_PLAIN waits per insertion)");
    std::vector<int> todo;
    const int rc = run_plan(c, "er_run", todo);
    if (rc) return rc;
    return run_slots(c, todo, mode);
}

// er_run without allocation: every run buffer must already hold this run's
// shapes, and the partial sums go to the scratch array; no events are
// recorded (profiling is suspended for the call)
extern "C" int er_run_noalloc(er_ctx *c, int32_t mode)
{
    if (!c) return ER_E_ARG;
    if (mode != ER_PLAIN && mode != ER_BATCHED && mode != ER_PLAIN_ASYNC)
        return fail(c, ER_E_ARG, "er_run_noalloc: mode");
    if (c->capturing) return fail(c, ER_E_STATE, "er_run_noalloc: not while a graph is recorded");
    if (!c->scratch_dev) return fail(c, ER_E_STATE, "er_run_noalloc: no scratch address for this device");
    std::vector<int> todo;
    int rc = run_plan(c, "er_run_noalloc", todo);
    if (rc) return rc;
    const bool prof = c->prof;
    c->prof = false;
    c->noalloc = true;
    rc = run_slots(c, todo, mode);
    c->noalloc = false;
    c->prof = prof;
    return rc;
}

// the halt (or, warn = true, the 'mbin' warning) of slot c->S():
// (insertion, molecule) from the key, the energy as captured at the end of
// the run that produced it
static void decode_halt(er_ctx *c, unsigned long long key, bool warn)
{
    const int jp1 = (int)(key & 0x1fffffu);
    unsigned long long cap[4] = {~0ull, 0ull, ~0ull, 0ull};
    (void)hipMemcpy(cap, c->S().d_halt.p, sizeof cap, hipMemcpyDeviceToHost);
    const int o = warn ? 2 : 0;
    float v = 0.f;
    if (cap[o] == key) {
        const unsigned int bits = (unsigned int)cap[o + 1];
        std::memcpy(&v, &bits, 4);
    } else {
        v = NAN;  // not captured (cannot happen after a completed run)
    }
    c->halt_energy = v;
    if (warn) {
        std::strcpy(c->halt_code, "mbin");
        c->halt_species = jp1 == 0 ? 0 : c->uvspec[jp1 - 1];
        return;
    }
    if (jp1 == 0) {
        std::strcpy(c->halt_code, "eng_min");
        c->halt_species = 0;
        return;
    }
    const int sp = c->uvspec[jp1 - 1];
    c->halt_species = sp;
    int lo = 0, hi = c->uvmax[sp - 1];
    const double *mesh = c->uvcrd.data() + c->idmin[sp - 1];
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if ((double)v > mesh[mid]) lo = mid + 1;
        else hi = mid;
    }
    std::strcpy(c->halt_code, lo <= 0 ? "eng_min" : "eng_sft");
}

extern "C" int er_sync(er_ctx *c)
{
    if (!c) return ER_E_ARG;
    ER_NO_CAPTURE(c, "er_sync");
    HIPCHK(c, hipStreamSynchronize(c->stream));
    const int keep = c->cur;
    for (int k = 0; k < (int)c->slots.size(); ++k) {
        unsigned long long key = ~0ull, wkey = ~0ull;
        HIPCHK(c, hipMemcpy(&key, c->slots[k].d_errkey.p, sizeof key, hipMemcpyDeviceToHost));
        HIPCHK(c, hipMemcpy(&wkey, c->slots[k].d_warnkey.p, sizeof wkey, hipMemcpyDeviceToHost));
        // getiduv tests the halt conditions before the warning, so at equal
        // keys (one energy) the halt is met first
        const bool warn = !c->prm.force_calculation && wkey < key;
        if (warn) key = wkey;
        if (key != ~0ull) {
            c->cur = k;
            decode_halt(c, key, warn);
            c->cur = keep;
            c->halt_slot = k;
            if (warn)
                return fail(c, ER_E_HALT,
                            "warning mbin: the maximum binning energy is too low for this species (solute %d, "
                            "species %d, energy %g); force_calculation lets the run go on",
                            k + 1, c->halt_species, (double)c->halt_energy);
            return fail(c, ER_E_HALT, "halt condition %s (solute %d, species %d, energy %g)", c->halt_code, k + 1,
                        c->halt_species, (double)c->halt_energy);
        }
    }
    return ER_OK;
}

extern "C" int er_get_uvengy(er_ctx *c, float *uvengy, float *uvengy0)
{
    if (!c) return ER_E_ARG;
    ER_NO_CAPTURE(c, "er_get_uvengy");
    if (c->S().last_nins <= 0) return fail(c, ER_E_STATE, "er_get_uvengy: no run yet");
    HIPCHK(c, hipStreamSynchronize(c->stream));
    const int nins = c->S().last_nins, NI = c->S().last_NI, nmol = c->nummol;
    if (uvengy) {
        const size_t ntile = (size_t)(nins + NI - 1) / NI;
        std::vector<float> u(ntile * NI * nmol);
        HIPCHK(c, hipMemcpy(u.data(), c->S().U_view, u.size() * sizeof(float), hipMemcpyDeviceToHost));
        for (int n = 0; n < nins; ++n)
            for (int j = 0; j < nmol; ++j) uvengy[(size_t)n * nmol + j] = u[((size_t)(n / NI) * nmol + j) * NI + n % NI];
    }
    if (uvengy0) HIPCHK(c, hipMemcpy(uvengy0, c->S().u0_view, nins * sizeof(float), hipMemcpyDeviceToHost));
    return ER_OK;
}

extern "C" int er_get_nins(er_ctx *c, int32_t *nins)
{
    if (!c || !nins) return ER_E_ARG;
    *nins = std::max(0, c->S().last_nins);
    return ER_OK;
}

extern "C" int er_fetch(er_ctx *c, double *edens, double *ecorr, double *eself, double *scal, float *minuv,
                        float *maxuv, double *slnuv)
{
    if (!c) return ER_E_ARG;
    ER_NO_CAPTURE(c, "er_fetch");
    if (!c->have_mesh) return fail(c, ER_E_STATE, "er_fetch: no mesh");
    HIPCHK(c, hipStreamSynchronize(c->stream));
    if (edens) HIPCHK(c, hipMemcpy(edens, c->S().d_edens.p, c->ermax * sizeof(double), hipMemcpyDeviceToHost));
    if (ecorr)
        HIPCHK(c, hipMemcpy(ecorr, c->S().d_ecorr.p, (size_t)c->ermax * c->ermax * sizeof(double), hipMemcpyDeviceToHost));
    if (eself && c->esmax > 0) HIPCHK(c, hipMemcpy(eself, c->S().d_eself.p, c->esmax * sizeof(double), hipMemcpyDeviceToHost));
    if (scal) HIPCHK(c, hipMemcpy(scal, c->S().d_state.p, 4 * sizeof(double), hipMemcpyDeviceToHost));
    if (minuv || maxuv) {
        std::vector<unsigned int> mm(2 * (size_t)(c->numslv + 1));
        HIPCHK(c, hipMemcpy(mm.data(), c->S().d_minmax.p, mm.size() * sizeof(unsigned int), hipMemcpyDeviceToHost));
        auto unord = [](unsigned int u) {
            u = (u & 0x80000000u) ? (u & 0x7fffffffu) : ~u;
            float f;
            std::memcpy(&f, &u, 4);
            return f;
        };
        for (int k = 0; k <= c->numslv; ++k) {
            if (minuv) minuv[k] = unord(mm[2 * k]);
            if (maxuv) maxuv[k] = unord(mm[2 * k + 1]);
        }
    }
    if (slnuv) HIPCHK(c, hipMemcpy(slnuv, c->S().d_slnuv.p, c->numslv * sizeof(double), hipMemcpyDeviceToHost));
    return ER_OK;
}

extern "C" int er_get_flceng(er_ctx *c, double *flceng)
{
    if (!c || !flceng) return ER_E_ARG;
    ER_NO_CAPTURE(c, "er_get_flceng");
    if (c->prm.slttype != 1 || c->S().last_nins <= 0) return fail(c, ER_E_STATE, "er_get_flceng: no soln run");
    HIPCHK(c, hipStreamSynchronize(c->stream));
    HIPCHK(c, hipMemcpy(flceng, c->d_flceng.p, (size_t)c->S().last_nins * c->numslv * sizeof(double), hipMemcpyDeviceToHost));
    return ER_OK;
}

extern "C" int er_clear(er_ctx *c)
{
    if (!c) return ER_E_ARG;
    ER_NO_CAPTURE(c, "er_clear");
    if (!c->have_mesh) return fail(c, ER_E_STATE, "er_clear: no mesh");
    for (Slot &sl : c->slots) {
        const int rc = slot_clear(c, sl);
        if (rc) return rc;
    }
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return ER_OK;
}

extern "C" int er_last_error(er_ctx *c, char *code, int32_t *species, float *energy)
{
    if (!c) return ER_E_ARG;
    if (code) std::memcpy(code, c->halt_code, 8);
    if (species) *species = c->halt_species;
    if (energy) *energy = c->halt_energy;
    return ER_OK;
}

extern "C" const char *er_error_string(er_ctx *c)
{
    if (!c) return "null context";
    return c->errmsg.c_str();
}

extern "C" int er_profile(er_ctx *c, int32_t enable)
{
    if (!c) return ER_E_ARG;
    ER_NO_CAPTURE(c, "er_profile");
    prof_collect(c);
    c->prof = enable != 0;
    return ER_OK;
}

extern "C" int er_profile_read(er_ctx *c, int32_t kernel, double *total_ms, int64_t *launches)
{
    if (!c) return ER_E_ARG;
    ER_NO_CAPTURE(c, "er_profile_read");
    if (kernel < 0 || kernel >= ER_K_COUNT) return fail(c, ER_E_ARG, "er_profile_read: kernel id");
    prof_collect(c);
    if (total_ms) *total_ms = c->prof_ms[kernel];
    if (launches) *launches = c->prof_n[kernel];
    c->prof_ms[kernel] = 0.0;
    c->prof_n[kernel] = 0;
    return ER_OK;
}

// ---- launch-shape tuning on the context's own frame and insertions ----

// the per-slot device state a tuning run changes (block accumulators,
// voffset, flags, minuv/maxuv, halt keys), kept aside while tuning
struct SlotSnap {
    DBuf<double> edens, ecorr, eself, state, slnuv;
    DBuf<int> flags;
    DBuf<float> fstate;
    DBuf<unsigned int> minmax;
    DBuf<unsigned long long> errkey, warnkey, halt;
};

template <class T> static hipError_t snap_take(DBuf<T> &dst, const DBuf<T> &src, hipStream_t s)
{
    if (!src.p || src.n == 0) return hipSuccess;
    hipError_t e = dst.ensure(src.n);
    if (e != hipSuccess) return e;
    return hipMemcpyAsync(dst.p, src.p, src.n * sizeof(T), hipMemcpyDeviceToDevice, s);
}

template <class T> static hipError_t snap_give(DBuf<T> &dst, const DBuf<T> &src, hipStream_t s)
{
    if (!src.p || src.n == 0) return hipSuccess;
    return hipMemcpyAsync(dst.p, src.p, src.n * sizeof(T), hipMemcpyDeviceToDevice, s);
}

static hipError_t slot_snap(Slot &sl, SlotSnap &sn, hipStream_t s, bool take)
{
    auto f = [&](auto &live, auto &kept) { return take ? snap_take(kept, live, s) : snap_give(live, kept, s); };
    hipError_t e;
    if ((e = f(sl.d_edens, sn.edens)) != hipSuccess || (e = f(sl.d_ecorr, sn.ecorr)) != hipSuccess ||
        (e = f(sl.d_eself, sn.eself)) != hipSuccess || (e = f(sl.d_state, sn.state)) != hipSuccess ||
        (e = f(sl.d_slnuv, sn.slnuv)) != hipSuccess || (e = f(sl.d_flags, sn.flags)) != hipSuccess ||
        (e = f(sl.d_warnkey, sn.warnkey)) != hipSuccess || (e = f(sl.d_fstate, sn.fstate)) != hipSuccess ||
        (e = f(sl.d_minmax, sn.minmax)) != hipSuccess || (e = f(sl.d_errkey, sn.errkey)) != hipSuccess ||
        (e = f(sl.d_halt, sn.halt)) != hipSuccess)
        return e;
    return hipSuccess;
}

struct EventPair {
    hipEvent_t a = nullptr, b = nullptr;
    ~EventPair()
    {
        if (a) (void)hipEventDestroy(a);
        if (b) (void)hipEventDestroy(b);
    }
};

extern "C" int er_set_cell_width(er_ctx *c, float angstrom);

// a candidate replaces the current shape only when it is faster by more
// than this fraction, so that timing noise keeps the current one
constexpr double TUNE_MARGIN = 0.02;

// the tuning runs proper: chunk over whole batched runs of every slot, then
// each slot's pair split from the pair kernel's own event times.  The
// current shapes are always among the candidates.
static int autotune_runs(er_ctx *c, const std::vector<int> &todo, int reps)
{
    // the run's insertions, every solute end to end (each padded to NI_B)
    int ntot = 0;
    for (int k : todo) ntot += ((c->prm.slttype == 1 ? c->nsltmol : c->slots[k].nins) + NI_B - 1) / NI_B * NI_B;
    // plane spectra and potentials of a chunk; candidates beyond 40 GiB of
    // them are skipped unless the current chunk already takes more.  The
    // histogram's slicing follows the chunk, so a chunk beyond the
    // insertions still differs from a smaller one; past the larger of the
    // current chunk and the run's insertions nothing changes.  The whole run
    // in one chunk is a candidate (one launch per kernel class).
    const int cur_chunk = c->chunk;
    const size_t per_ins = c->Gc * sizeof(float2) + c->G * sizeof(float) + (size_t)c->numatm * sizeof(float);
    const size_t cap = std::max<size_t>((size_t)40 << 30, (size_t)cur_chunk * per_ins);
    const int reach = std::max(cur_chunk, ntot);
    std::vector<int> chunks = {cur_chunk};
    for (int ch : {256, 512, 1024, 2048, 4096, 8192, 16384, ntot})
        if (ch != cur_chunk && ch <= reach && (size_t)ch * per_ins <= cap &&
            std::find(chunks.begin(), chunks.end(), ch) == chunks.end())
            chunks.push_back(ch);
    EventPair ev;
    HIPCHK(c, hipEventCreate(&ev.a));
    HIPCHK(c, hipEventCreate(&ev.b));
    int rc;
    if (chunks.size() > 1) {
        int best = cur_chunk;
        float best_ms = 0.f;
        for (size_t i = 0; i < chunks.size(); ++i) {
            c->chunk = chunks[i];
            if ((rc = run_slots(c, todo, ER_BATCHED))) return rc;  // sizes the buffers
            HIPCHK(c, hipEventRecord(ev.a, c->stream));
            for (int r = 0; r < reps; ++r)
                if ((rc = run_slots(c, todo, ER_BATCHED))) return rc;
            HIPCHK(c, hipEventRecord(ev.b, c->stream));
            HIPCHK(c, hipEventSynchronize(ev.b));
            float ms = 0.f;
            HIPCHK(c, hipEventElapsedTime(&ms, ev.a, ev.b));
            if (i == 0 || ms < best_ms * (1.0 - TUNE_MARGIN)) {
                best_ms = ms;
                best = chunks[i];
            }
        }
        c->chunk = best;
    }
    c->prof = true;
    // histogram slices (refs runs): the histogram kernel's own time over
    // whole runs; slices change only integer counts and minima/maxima
    if (c->prm.slttype != 1) {
        int best = c->hist_slices;
        double best_ms = 0.0;
        std::vector<int> nsl = {c->hist_slices};
        for (int v : {1, 2, 4, 8, 16, 32, 64})
            if (v != c->hist_slices) nsl.push_back(v);
        for (size_t i = 0; i < nsl.size(); ++i) {
            c->hist_slices = nsl[i];
            if ((rc = run_slots(c, todo, ER_BATCHED))) return rc;
            prof_collect(c);
            c->prof_ms[ER_K_HIST] = 0.0;
            for (int r = 0; r < reps; ++r)
                if ((rc = run_slots(c, todo, ER_BATCHED))) return rc;
            prof_collect(c);
            if (i == 0 || c->prof_ms[ER_K_HIST] < best_ms * (1.0 - TUNE_MARGIN)) {
                best_ms = c->prof_ms[ER_K_HIST];
                best = nsl[i];
            }
        }
        c->hist_slices = best;
    }
    if (c->prm.boxshp == 0) return ER_OK;  // the open-system pair kernel has no split or cell list
    // the cell list's least cell width: the pair kernel's own time over whole
    // runs (the cell list is rebuilt for each candidate; cells partition the
    // molecules whatever their size, so no result changes)
    {
        const float keep_w = c->cell_min;
        float best = keep_w;
        double best_ms = 0.0;
        std::vector<float> ws = {keep_w};
        for (float w : {3.5f, 5.0f, 7.0f, 10.0f})
            if (w != keep_w && !(keep_w == 0.f && w == 3.5f)) ws.push_back(w);
        for (size_t i = 0; i < ws.size(); ++i) {
            if ((rc = er_set_cell_width(c, ws[i]))) return rc;
            if ((rc = run_slots(c, todo, ER_BATCHED))) return rc;
            prof_collect(c);
            c->prof_ms[ER_K_PAIRS] = 0.0;
            for (int r = 0; r < reps; ++r) {
                c->frame_dirty = true;  // each run rebuilds the cell list, as a new frame would
                if ((rc = run_slots(c, todo, ER_BATCHED))) return rc;
            }
            prof_collect(c);
            if (i == 0 || c->prof_ms[ER_K_PAIRS] < best_ms * (1.0 - TUNE_MARGIN)) {
                best_ms = c->prof_ms[ER_K_PAIRS];
                best = ws[i];
            }
        }
        if ((rc = er_set_cell_width(c, best))) return rc;
    }
    for (int k : todo) {
        c->cur = k;
        int best = c->S().pair_split;
        double best_ms = 0.0;
        std::vector<int> splits = {c->S().pair_split};  // 0: the built-in choice
        for (int sp : {1, 2, 4, 8, 16, 32, 64})
            if (sp != c->S().pair_split) splits.push_back(sp);
        for (size_t i = 0; i < splits.size(); ++i) {
            c->S().pair_split = splits[i];
            const std::vector<int> one = {k};
            if ((rc = run_slots(c, one, ER_BATCHED))) return rc;
            prof_collect(c);
            c->prof_ms[ER_K_PAIRS] = 0.0;
            for (int r = 0; r < reps; ++r)
                if ((rc = run_slots(c, one, ER_BATCHED))) return rc;
            prof_collect(c);
            if (i == 0 || c->prof_ms[ER_K_PAIRS] < best_ms * (1.0 - TUNE_MARGIN)) {
                best_ms = c->prof_ms[ER_K_PAIRS];
                best = splits[i];
            }
        }
        c->S().pair_split = best;
    }
    return ER_OK;
}

extern "C" int er_autotune(er_ctx *c, int32_t reps)
{
    if (!c) return ER_E_ARG;
    ER_NO_CAPTURE(c, "er_autotune");
    if (reps < 0 || reps > 100) return fail(c, ER_E_ARG, "er_autotune: reps must lie in 0..100");
    if (reps == 0) reps = 3;
    std::vector<int> todo;
    int rc = run_plan(c, "er_autotune", todo);
    if (rc) return rc;
    // the caller's profile totals and every slot's accumulators are kept
    prof_collect(c);
    const bool prof_keep = c->prof;
    double ms_keep[ER_K_COUNT];
    int64_t n_keep[ER_K_COUNT];
    std::memcpy(ms_keep, c->prof_ms, sizeof ms_keep);
    std::memcpy(n_keep, c->prof_n, sizeof n_keep);
    std::vector<SlotSnap> snap(c->slots.size());
    for (size_t k = 0; k < c->slots.size(); ++k) HIPCHK(c, slot_snap(c->slots[k], snap[k], c->stream, true));
    const int keep = c->cur;
    // the shapes in force, restored if tuning fails part way
    const int chunk_keep = c->chunk, slices_keep = c->hist_slices;
    const float cellw_keep = c->cell_min;
    std::vector<int> split_keep(c->slots.size());
    for (size_t k = 0; k < c->slots.size(); ++k) split_keep[k] = c->slots[k].pair_split;
    rc = autotune_runs(c, todo, reps);
    c->cur = keep;
    if (rc == ER_OK) rc = er_sync(c);  // a halt hit by a tuning run is reported here
    if (rc != ER_OK) {
        c->chunk = chunk_keep;
        c->hist_slices = slices_keep;
        (void)er_set_cell_width(c, cellw_keep);
        for (size_t k = 0; k < c->slots.size(); ++k) c->slots[k].pair_split = split_keep[k];
    }
    (void)hipStreamSynchronize(c->stream);
    // spectra and potentials sized for a larger candidate chunk than the one
    // kept are given back
    {
        const size_t ntile = (size_t)(c->chunk + NI_B - 1) / NI_B;
        ++c->alloc_gen;
        if (c->d_T.n > (size_t)c->chunk * c->Gc) c->d_T.release();
        if (c->d_phi.n > ntile * NI_B * c->G) c->d_phi.release();
        if (c->d_Eat.n > ntile * NI_B * (size_t)c->numatm) c->d_Eat.release();
    }
    // one run with the kept shapes sizes the buffers and the device
    // descriptors for them, so a graph can be recorded right after tuning
    // (its accumulation is undone below with the tuning runs')
    if (rc == ER_OK) rc = run_slots(c, todo, ER_BATCHED);
    (void)hipStreamSynchronize(c->stream);
    prof_collect(c);
    std::memcpy(c->prof_ms, ms_keep, sizeof ms_keep);
    std::memcpy(c->prof_n, n_keep, sizeof n_keep);
    c->prof = prof_keep;
    for (size_t k = 0; k < c->slots.size(); ++k) HIPCHK(c, slot_snap(c->slots[k], snap[k], c->stream, false));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return rc;
}

extern "C" int er_set_tuning(er_ctx *c, int32_t chunk, int32_t pair_split, int32_t hist_slices)
{
    if (!c) return ER_E_ARG;
    ER_NO_CAPTURE(c, "er_set_tuning");
    if (chunk != 0 && (chunk < NI_B || chunk > 65536 || chunk % NI_B != 0))
        return fail(c, ER_E_ARG, "er_set_tuning: chunk must be 0 or a multiple of %d in %d..65536", NI_B, NI_B);
    if (pair_split < 0 || pair_split > 64) return fail(c, ER_E_ARG, "er_set_tuning: pair_split must lie in 0..64");
    if (hist_slices < 0 || hist_slices > 64) return fail(c, ER_E_ARG, "er_set_tuning: hist_slices must lie in 0..64");
    if (chunk != 0) c->chunk = chunk;
    c->S().pair_split = pair_split;
    c->hist_slices = hist_slices;
    return ER_OK;
}

extern "C" int er_get_tuning(er_ctx *c, int32_t *chunk, int32_t *pair_split, int32_t *hist_slices)
{
    if (!c) return ER_E_ARG;
    if (chunk) *chunk = c->chunk;
    if (pair_split) *pair_split = c->S().pair_split;
    if (hist_slices) *hist_slices = c->hist_slices;
    return ER_OK;
}

// ---- HIP graphs of whole runs ----
//
// er_graph_begin .. er_graph_end records the context's work (frames set on
// the device, their insertions, batched runs) into one HIP graph that
// er_graph_launch replays with a single launch.  The recorded kernels keep
// the device pointers, cells, weights and launch shapes of the recording;
// the caller refills the same device buffers between replays.  The host-side
// state a replay leaves behind (frame, cell, per-solute energies) is the one
// the recording left, restored at each launch.

struct HostRunState {
    bool have_frame, frame_dirty, cell_changed, have_prev_cell;
    float cell[9], prev_cell[9];
    FrameConst fc;
    double invcl_d[9], volume_d;
    float w_sys;
    const float *d_sitepos;
    int nc, ncell;
    float cell_w;
    struct SlotState {
        bool have_ins, wslt_given;
        int nins, last_NI, last_nins;
        const float *d_sltpos, *U_view, *u0_view;
    };
    std::vector<SlotState> slots;
};

struct er_ctx::GraphRec {
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    unsigned long long gen = 0;
    std::vector<SlotDesc> sd;
    bool engfac_at_start = false;  // the graph computes the influence function before its first run
    bool have_prev_begin = false;
    float prev_begin[9] = {0};
    HostRunState end;
    size_t nslots = 0;
};

static void save_state(const er_ctx *c, HostRunState &h)
{
    h.have_frame = c->have_frame;
    h.frame_dirty = c->frame_dirty;
    h.cell_changed = c->cell_changed;
    h.have_prev_cell = c->have_prev_cell;
    std::memcpy(h.cell, c->cell, sizeof h.cell);
    std::memcpy(h.prev_cell, c->prev_cell, sizeof h.prev_cell);
    h.fc = c->fc;
    std::memcpy(h.invcl_d, c->invcl_d, sizeof h.invcl_d);
    h.volume_d = c->volume_d;
    h.w_sys = c->w_sys;
    h.d_sitepos = c->d_sitepos;
    h.nc = c->nc;
    h.ncell = c->ncell;
    h.cell_w = c->cell_w;
    h.slots.resize(c->slots.size());
    for (size_t k = 0; k < c->slots.size(); ++k) {
        const Slot &sl = c->slots[k];
        h.slots[k] = {sl.have_ins, sl.wslt_given, sl.nins, sl.last_NI, sl.last_nins, sl.d_sltpos, sl.U_view,
                      sl.u0_view};
    }
}

static void restore_state(er_ctx *c, const HostRunState &h)
{
    c->have_frame = h.have_frame;
    c->frame_dirty = h.frame_dirty;
    c->cell_changed = h.cell_changed;
    c->have_prev_cell = h.have_prev_cell;
    std::memcpy(c->cell, h.cell, sizeof h.cell);
    std::memcpy(c->prev_cell, h.prev_cell, sizeof h.prev_cell);
    c->fc = h.fc;
    std::memcpy(c->invcl_d, h.invcl_d, sizeof h.invcl_d);
    c->volume_d = h.volume_d;
    c->w_sys = h.w_sys;
    c->d_sitepos = h.d_sitepos;
    c->nc = h.nc;
    c->ncell = h.ncell;
    c->cell_w = h.cell_w;
    for (size_t k = 0; k < c->slots.size() && k < h.slots.size(); ++k) {
        Slot &sl = c->slots[k];
        const auto &s = h.slots[k];
        sl.have_ins = s.have_ins;
        sl.wslt_given = s.wslt_given;
        sl.nins = s.nins;
        sl.last_NI = s.last_NI;
        sl.last_nins = s.last_nins;
        sl.d_sltpos = s.d_sltpos;
        sl.U_view = s.U_view;
        sl.u0_view = s.u0_view;
    }
}

static void graph_free(er_ctx::GraphRec *g)
{
    if (!g) return;
    if (g->exec) (void)hipGraphExecDestroy(g->exec);
    if (g->graph) (void)hipGraphDestroy(g->graph);
    delete g;
}

extern "C" int er_graph_begin(er_ctx *c)
{
    if (!c) return ER_E_ARG;
    ER_NO_CAPTURE(c, "er_graph_begin");
    if (!c->have_mesh) return fail(c, ER_E_STATE, "er_graph_begin: er_set_mesh first");
    // the frame work of an earlier eager run has joined the main stream
    HIPCHK(c, join_atoms(c));
    c->prof_before_capture = c->prof;
    c->prof = false;
    c->cap_have_prev = c->have_prev_cell;
    std::memcpy(c->cap_prev_cell, c->prev_cell, sizeof c->cap_prev_cell);
    c->cap_gen = c->alloc_gen;
    c->capture_bad = false;
    c->capture_engfac = false;
    c->capture_runs = 0;
    const hipError_t e = hipStreamBeginCapture(c->stream, hipStreamCaptureModeRelaxed);
    if (e != hipSuccess) {
        c->prof = c->prof_before_capture;
        return fail(c, ER_E_HIP, "er_graph_begin: hipStreamBeginCapture: %s", hipGetErrorString(e));
    }
    c->capturing = true;
    return ER_OK;
}

extern "C" int er_graph_end(er_ctx *c, int32_t *graph)
{
    if (!c || !graph) return ER_E_ARG;
    *graph = 0;
    if (!c->capturing) return fail(c, ER_E_STATE, "er_graph_end: no graph is being recorded");
    hipGraph_t g = nullptr;
    const hipError_t e = hipStreamEndCapture(c->stream, &g);
    c->capturing = false;
    c->prof = c->prof_before_capture;
    auto drop = [&](int code, const char *why) {
        if (g) (void)hipGraphDestroy(g);
        return fail(c, code, "er_graph_end: %s", why);
    };
    if (e != hipSuccess) return drop(ER_E_HIP, hipGetErrorString(e));
    if (c->capture_bad) return drop(ER_E_STATE, "a call failed while recording; the graph is discarded");
    if (c->capture_runs == 0) return drop(ER_E_STATE, "no er_run was recorded");
    if (c->alloc_gen != c->cap_gen) return drop(ER_E_STATE, "device buffers changed while recording");
    auto *r = new er_ctx::GraphRec();
    r->graph = g;
    const hipError_t ei = hipGraphInstantiate(&r->exec, g, nullptr, nullptr, 0);
    if (ei != hipSuccess) {
        graph_free(r);
        return fail(c, ER_E_HIP, "er_graph_end: hipGraphInstantiate: %s", hipGetErrorString(ei));
    }
    r->gen = c->alloc_gen;
    r->sd = c->sd_up;
    r->engfac_at_start = c->capture_engfac;
    r->have_prev_begin = c->cap_have_prev;
    std::memcpy(r->prev_begin, c->cap_prev_cell, sizeof r->prev_begin);
    save_state(c, r->end);
    r->nslots = c->slots.size();
    // the recording launched nothing: the state it leaves is what a replay leaves
    c->graphs.push_back(r);
    *graph = (int32_t)c->graphs.size();
    return ER_OK;
}

extern "C" int er_graph_launch(er_ctx *c, int32_t graph)
{
    if (!c) return ER_E_ARG;
    ER_NO_CAPTURE(c, "er_graph_launch");
    if (graph < 1 || graph > (int32_t)c->graphs.size() || !c->graphs[graph - 1])
        return fail(c, ER_E_ARG, "er_graph_launch: no graph %d", graph);
    er_ctx::GraphRec *r = c->graphs[graph - 1];
    if (r->gen != c->alloc_gen || r->nslots != c->slots.size())
        return fail(c, ER_E_STATE, "er_graph_launch: device buffers were reallocated since graph %d was recorded; "
                                   "record it again", graph);
    if (r->sd.size() != c->sd_up.size() ||
        std::memcmp(r->sd.data(), c->sd_up.data(), r->sd.size() * sizeof(SlotDesc)) != 0)
        return fail(c, ER_E_STATE, "er_graph_launch: the solutes or launch shapes changed since graph %d was "
                                   "recorded; record it again", graph);
    if (!r->engfac_at_start &&
        (c->have_prev_cell != r->have_prev_begin ||
         (r->have_prev_begin && std::memcmp(c->prev_cell, r->prev_begin, sizeof r->prev_begin) != 0)))
        return fail(c, ER_E_STATE, "er_graph_launch: graph %d was recorded for another cell's influence function",
                    graph);
    HIPCHK(c, join_atoms(c));
    HIPCHK(c, hipGraphLaunch(r->exec, c->stream));
    restore_state(c, r->end);
    return ER_OK;
}

extern "C" int er_graph_destroy(er_ctx *c, int32_t graph)
{
    if (!c) return ER_E_ARG;
    ER_NO_CAPTURE(c, "er_graph_destroy");
    if (graph < 1 || graph > (int32_t)c->graphs.size() || !c->graphs[graph - 1])
        return fail(c, ER_E_ARG, "er_graph_destroy: no graph %d", graph);
    HIPCHK(c, hipStreamSynchronize(c->stream));
    graph_free(c->graphs[graph - 1]);
    c->graphs[graph - 1] = nullptr;
    return ER_OK;
}

// the least cell width of the pair search's cell list: the cells of the next
// run are rebuilt for it (the frame's cell is kept)
extern "C" int er_set_cell_width(er_ctx *c, float angstrom)
{
    if (!c) return ER_E_ARG;
    ER_NO_CAPTURE(c, "er_set_cell_width");
    if (!(angstrom == 0.f || (angstrom >= 1.0f && angstrom <= 1000.0f)))
        return fail(c, ER_E_ARG, "er_set_cell_width: 0 (built-in) or 1..1000 Angstrom");
    c->cell_min = angstrom;
    if (c->have_frame && c->prm.boxshp != 0 && c->cell_wmin > 0.0) {
        c->nc = std::max(1, std::min(32, (int)(c->cell_wmin / (c->cell_min > 0.f ? c->cell_min : 3.5f))));
        c->ncell = morton_keys(c->nc);
        c->cell_w = (float)(c->cell_wmin / c->nc);
        c->frame_dirty = true;
    }
    return ER_OK;
}

extern "C" int er_get_cell_width(er_ctx *c, float *angstrom)
{
    if (!c || !angstrom) return ER_E_ARG;
    *angstrom = c->cell_min;
    return ER_OK;
}

extern "C" void er_destroy(er_ctx *c)
{
    if (!c) return;
    if (c->capturing) {  // a recording left open is discarded
        hipGraph_t g = nullptr;
        (void)hipStreamEndCapture(c->stream, &g);
        if (g) (void)hipGraphDestroy(g);
        c->capturing = false;
    }
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    prof_collect(c);
    for (hipEvent_t e : c->prof_pool) (void)hipEventDestroy(e);
    c->prof_pool.clear();
    c->d_charge.release(); c->d_ljene.release(); c->d_ljlensq.release(); c->d_molq.release();
    c->d_ljtype.release(); c->d_molbeg.release(); c->d_spb.release(); c->d_spe.release();
    for (Slot &sl : c->slots) sl.release();
    c->d_tag.release();
    c->d_uvcrd.release(); c->d_escrd.release(); c->d_uvmax.release(); c->d_uvsoft.release(); c->d_idmin.release();
    c->d_sitepos_own.release(); c->d_posn.release(); c->d_posu.release(); c->d_molref.release(); c->d_segref.release(); c->d_splw.release(); c->d_splg.release();
    c->d_insposn.release(); c->d_insposu.release(); c->d_insref.release(); c->d_insw.release(); c->d_selfreal.release(); c->d_insg.release();
    c->d_splfc1.release(); c->d_splfc2.release(); c->d_splfc3.release(); c->d_engfac.release();
    c->d_tw1.release(); c->d_tw2.release(); c->d_tw3.release(); c->d_twm.release(); c->d_tws.release();
    c->d_selfpart.release(); c->d_Ftab.release(); c->d_T.release(); c->d_phi.release();
    c->d_C.release(); c->d_w.release(); c->d_border.release(); c->d_bdesc.release(); c->d_flceng.release(); c->d_part.release(); c->d_epart.release();
    if (c->stream2) (void)hipStreamSynchronize(c->stream2);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    if (c->stream2) (void)hipStreamDestroy(c->stream2);
    if (c->ev_in) (void)hipEventDestroy(c->ev_in);
    if (c->ev_atoms) (void)hipEventDestroy(c->ev_atoms);
    for (auto *g : c->graphs) graph_free(g);
    delete c;
}
