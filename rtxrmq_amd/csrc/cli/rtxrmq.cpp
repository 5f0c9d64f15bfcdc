// rtxrmq.cpp -- the ./rtxrmq <n> <q> <lr> <alg> command line over librtxrmq.so.
//
// Same positional arguments, options, algorithm ids and CSV rows as the
// reference driver (src/main.cu:51-171, src/tools.h:28-189,268-336), so its
// scripts/*.sh sweeps and plotting scripts run unchanged.  The RTX-core algs
// (3, 4, 6, 8, 9, 10, 103) are not part of this engine; alg 5 / 105 run the
// HIP block engine, 2 / 102 the HIP brute force, 0 / 100 the host brute force,
// 1 / 101 the host HRMQ (rmq_hrmq_*, rightmost minimum on ties).
#include <getopt.h>
#include <hip/hip_runtime.h>
#include <unistd.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <ctime>
#include <string>
#include <thread>
#include <vector>

#include "power.h"
#include "rmq.h"

namespace {

#define C_RESET "\033[0m"
#define C_GREEN "\033[32m"
#define C_YELLOW "\033[33m"
#define C_CYAN "\033[1m\033[36m"
#define C_RED "\033[1m\033[31m"

enum : int {
    ALG_CPU_BASE = 0, ALG_CPU_HRMQ = 1, ALG_GPU_BASE = 2, ALG_GPU_RTX_CAST = 3, ALG_GPU_RTX_TRANS = 4,
    ALG_GPU_RTX_BLOCKS = 5, ALG_GPU_RTX_LUP = 6, ALG_GPU_RTX_IAS = 8, ALG_GPU_RTX_IAS_TRANS = 9,
    ALG_GPU_RTX_SER = 10, ALG_CPU_BASE_IDX = 100, ALG_CPU_HRMQ_IDX = 101, ALG_GPU_BASE_IDX = 102,
    ALG_GPU_RTX_CAST_IDX = 103, ALG_GPU_RTX_BLOCKS_IDX = 105
};

// The reference's labels (src/main.cu:35), kept so CSV rows group the same way.
const char *kAlgName[11] = {"[CPU] BASE", "[CPU] HRMQ", "[GPU] BASE", "[GPU] RTX_cast", "[GPU] RTX_trans",
                            "[GPU] RTX_blocks", "[GPU] RTX_lup", "[GPU] LCA (TODO)", "[GPU] RTX_ias",
                            "[GPU] RTX_ias_trans", "[GPU] RTX_ser"};

const char *alg_name(int alg) {
    const int a = alg % 100;
    return (a >= 0 && a <= 10) ? kAlgName[a] : "[?] unknown";
}

struct Args {
    int n = 0, q = 0, lr = 0, alg = 0, bs = 1 << 15, nb = 0, reps = 10, dev = 0, nt = 1, seed = 0;
    bool check = false, save_time = false, save_power = false;
    std::string time_file, power_file;
    int short_max = 0, group_lanes = 0, long_lanes = 0;
    int array_kind = RMQ_GEN_UNIFORM;
    bool qgen_ref = false;  // the reference's own query stream (src/rand.cuh:116-201)
    bool edges = false;     // build edge tables (RMQ_BUILD_EDGES)
    bool mid = false;       // with edge tables, coarse rows for an LDS superset test (RMQ_BUILD_MID)
    bool tune = false;      // pick bs, edge tables and thresholds with rmq_tune
    bool bs_given = false;  // --bs or --nb on the command line: --tune keeps that block size
    std::string tune_cache;
};

void usage() {
    fprintf(stderr,
            "usage: ./rtxrmq <n> <q> <lr> <alg> [options]\n\n"
            "  n    number of float32 elements\n"
            "  q    number of range-minimum queries\n"
            "  lr   range length: >0 fixed; -1 uniform in [1,n-l]; -2 lognormal median 2^14;\n"
            "       -3 lognormal median 2^6; -4 / -5 lognormal n/2^8, n/2^15 (sigma 0.3); -6 r = n-1\n"
            "  alg  0 %s (values) | 100 same, indices\n"
            "       1 %s | 101 indices  (host, --nt threads; rightmost index on ties)\n"
            "       2 %s | 102 indices\n"
            "       5 %s | 105 indices  (the HIP block engine)\n\n"
            "options:\n"
            "  --bs <block size>     elements per block for alg 5/105 (default 2^15)\n"
            "  --nb <#blocks>        number of blocks, bs = n/nb (overrides --bs)\n"
            "  --reps <r>            query repetitions for the mean time (default 10)\n"
            "  --dev <id>            HIP device (default 0)\n"
            "  --nt <threads>        host threads (query generation, CPU algs)\n"
            "  --seed <s>            generator seed (default time(0))\n"
            "  --check               compare with the GPU brute force (alg 2/102), bit for bit\n"
            "  --save-time[=file]    append a CSV row (default ../results/data.csv)\n"
            "  --save-power[=file]   sample power during the query loop: the GPU's (amdgpu hwmon, 100 ms)\n"
            "                        for GPU algs, the CPU packages' (powercap RAPL, 500 ms) for CPU algs\n"
            "  --short-max <len>     ranges up to this length use the group scan (alg 5/105)\n"
            "  --group-lanes <g>     lanes per short query: 4, 8, 16 or 64\n"
            "  --long-lanes <g>      lanes per long query: 4, 8, 16 or 64\n"
            "  --edges               build edge tables: O(1) answers for ranges not inside a 32-element sub-block\n"
            "  --mid                 with --edges: coarse rows in LDS tried before the block table\n"
            "  --tune[=cache]        autotune bs, edge tables and thresholds on the first 2^24 queries\n"
            "                        (overrides --edges/--mid and the lane options; with --bs/--nb the block size\n"
            "                        stays and only edge tables and thresholds are tuned; optional pick cache)\n"
            "  --array <kind>        uniform | ties16 | desc | const (default uniform)\n"
            "  --qgen <gen>          ours (default: counter-based, thread-count independent) |\n"
            "                        ref (the reference's mt19937(seed*tid) stream; depends on --nt;\n"
            "                        lr -1..-5 as the reference defines them)\n",
            kAlgName[0], kAlgName[1], kAlgName[2], kAlgName[5]);
}

bool is_block_alg(int alg) {
    return alg == ALG_GPU_RTX_CAST || alg == ALG_GPU_RTX_TRANS || alg == ALG_GPU_RTX_BLOCKS ||
           alg == ALG_GPU_RTX_BLOCKS_IDX || alg == ALG_GPU_RTX_SER || alg == ALG_GPU_RTX_LUP ||
           alg == ALG_GPU_RTX_IAS || alg == ALG_GPU_RTX_IAS_TRANS;
}

// src/tools.h:76-189: positional n q lr alg, then long options.
bool parse(int argc, char **argv, Args &a) {
    if (argc < 5) return false;
    a.n = atoi(argv[1]);
    a.q = atoi(argv[2]);
    a.lr = atoi(argv[3]);
    a.alg = atoi(argv[4]);
    if (a.n <= 0 || a.q <= 0 || a.lr == 0) return false;
    if (a.lr > a.n) {
        fprintf(stderr, "Error: lr=%i > n=%i  (lr must be between '1' and 'n')\n", a.lr, a.n);
        exit(EXIT_FAILURE);
    }
    a.bs = std::min(a.n, 1 << 15);
    a.nb = a.n / a.bs;
    a.seed = (int)time(nullptr);
    enum { O_BS = 1, O_NB, O_REPS, O_DEV, O_NT, O_SEED, O_CHECK, O_TIME, O_POWER, O_SHORT, O_GROUP, O_ARRAY, O_LONG, O_QGEN,
           O_EDGES, O_TUNE, O_MID };
    static struct option opts[] = {
        {"bs", required_argument, nullptr, O_BS},         {"nb", required_argument, nullptr, O_NB},
        {"reps", required_argument, nullptr, O_REPS},     {"dev", required_argument, nullptr, O_DEV},
        {"nt", required_argument, nullptr, O_NT},         {"seed", required_argument, nullptr, O_SEED},
        {"check", no_argument, nullptr, O_CHECK},         {"save-time", optional_argument, nullptr, O_TIME},
        {"save-power", optional_argument, nullptr, O_POWER},
        {"short-max", required_argument, nullptr, O_SHORT}, {"group-lanes", required_argument, nullptr, O_GROUP},
        {"array", required_argument, nullptr, O_ARRAY},   {"long-lanes", required_argument, nullptr, O_LONG},
        {"qgen", required_argument, nullptr, O_QGEN},     {"edges", no_argument, nullptr, O_EDGES},
        {"tune", optional_argument, nullptr, O_TUNE},     {"mid", no_argument, nullptr, O_MID},
        {nullptr, 0, nullptr, 0}};
    int opt, idx;
    optind = 5;
    while ((opt = getopt_long(argc, argv, "", opts, &idx)) != -1) {
        switch (opt) {
        case O_BS: a.bs = std::max(1, std::min(a.n, atoi(optarg))); a.nb = a.n / a.bs; a.bs_given = true; break;
        case O_NB: a.nb = std::max(1, std::min(a.n, atoi(optarg))); a.bs = a.n / a.nb; a.bs_given = true; break;
        case O_REPS: a.reps = std::max(1, atoi(optarg)); break;
        case O_DEV: a.dev = atoi(optarg); break;
        case O_NT: a.nt = std::max(1, atoi(optarg)); break;
        case O_SEED: a.seed = atoi(optarg); break;
        case O_CHECK: a.check = true; break;
        case O_TIME: a.save_time = true; if (optarg) a.time_file = optarg; break;
        case O_POWER: a.save_power = true; if (optarg) a.power_file = optarg; break;
        case O_SHORT: a.short_max = atoi(optarg); break;
        case O_GROUP: a.group_lanes = atoi(optarg); break;
        case O_LONG: a.long_lanes = atoi(optarg); break;
        case O_EDGES: a.edges = true; break;
        case O_MID: a.mid = true; break;
        case O_TUNE: a.tune = true; if (optarg) a.tune_cache = optarg; break;
        case O_QGEN:
            if (!strcmp(optarg, "ref")) a.qgen_ref = true;
            else if (strcmp(optarg, "ours")) return false;
            break;
        case O_ARRAY:
            if (!strcmp(optarg, "uniform")) a.array_kind = RMQ_GEN_UNIFORM;
            else if (!strcmp(optarg, "ties16")) a.array_kind = RMQ_GEN_TIES16;
            else if (!strcmp(optarg, "desc")) a.array_kind = RMQ_GEN_DESC;
            else if (!strcmp(optarg, "const")) a.array_kind = RMQ_GEN_CONST;
            else return false;
            break;
        default: return false;
        }
    }
    if (!is_block_alg(a.alg)) {  // src/tools.h:163-173
        a.bs = 0;
        a.nb = 0;
    }
    return true;
}

[[noreturn]] void die(const char *what, int rc) {
    fprintf(stderr, C_RED "%s failed: %s (%d)" C_RESET "\n", what, rmq_strerror(rc), rc);
    exit(EXIT_FAILURE);
}

void hip_ok(hipError_t e, const char *what) {
    if (e != hipSuccess) {
        fprintf(stderr, C_RED "%s failed: %s" C_RESET "\n", what, hipGetErrorString(e));
        exit(EXIT_FAILURE);
    }
}

double now_ms() {
    return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

std::string device_label(const Args &a) {
    if (a.alg == ALG_CPU_BASE || a.alg == ALG_CPU_HRMQ || a.alg == ALG_CPU_BASE_IDX || a.alg == ALG_CPU_HRMQ_IDX) {
        char host[64] = {0};
        gethostname(host, sizeof(host) - 1);
        return std::string("CPU ") + host;  // src/tools.h:279-284
    }
    hipDeviceProp_t p;
    if (hipGetDeviceProperties(&p, a.dev) != hipSuccess) return "unknown";
    return p.name[0] ? p.name : p.gcnArchName;  // the marketing name needs libdrm's ids table
}

std::string csv_path(const Args &a) { return a.time_file.empty() ? "../results/data.csv" : a.time_file; }

// Row prefix, src/tools.h:268-297: device,alg,reps,n,bs,q,lr
void csv_prefix(const Args &a) {
    if (!a.save_time) return;
    FILE *f = fopen(csv_path(a).c_str(), "a");
    if (!f) return;
    fprintf(f, "%s,%s,%i,%i,%i,%i,%i", device_label(a).c_str(), alg_name(a.alg), a.reps, a.n, a.bs, a.q, a.lr);
    fclose(f);
}

// Row suffix, src/tools.h:317-336: t (s, all reps), q/s, ns/q, construction ms, out MB, temp MB
void csv_suffix(const Args &a, double time_ms, double construction_ms, double out_bytes, double temp_bytes) {
    if (!a.save_time) return;
    FILE *f = fopen(csv_path(a).c_str(), "a");
    if (!f) return;
    const double per = time_ms / a.reps;
    fprintf(f, ",%f,%f,%f,%f,%f,%f\n", time_ms / 1000.0, (double)a.q / (per / 1000.0), per * 1e6 / a.q,
            construction_ms, out_bytes / 1e6, temp_bytes / 1e6);
    fclose(f);
}

void print_rate(const char *label, const Args &a, double time_ms) {
    const double per_s = time_ms / (1000.0 * a.reps);
    printf(C_CYAN "%s done: %f secs (avg %f secs): [%.2f RMQs/sec, %f nsec/RMQ]" C_RESET "\n", label,
           time_ms / 1000.0, per_s, (double)a.q / per_s, per_s * 1e9 / a.q);
}

void print_device(int dev) {
    hipDeviceProp_t p;
    if (hipGetDeviceProperties(&p, dev) != hipSuccess) return;
    printf("Device Number: %d\n  Device name:                  %s (%s)\n", dev, p.name, p.gcnArchName);
    printf("  Memory:                       %f GB\n", p.totalGlobalMem / (1024.0 * 1024.0 * 1024.0));
    printf("  Compute units:                %d\n", p.multiProcessorCount);
    printf("  Memory clock / bus:           %d MHz / %d bits\n\n", p.memoryClockRate / 1000, p.memoryBusWidth);
}

// Host brute force, algs 0/100 (src/cpu_methods.h:7-77): scan from A[l], strict '<'.
void cpu_base(const Args &a, const float *A, const int32_t *Q, float *val, int32_t *idx) {
    auto work = [&](int64_t lo, int64_t hi) {
        for (int64_t j = lo; j < hi; ++j) {
            const int32_t l = Q[2 * j], r = Q[2 * j + 1];
            float m = A[l];
            int32_t mi = l;
            for (int32_t i = l; i <= r; ++i)
                if (A[i] < m) { m = A[i]; mi = i; }
            if (val) val[j] = m;
            if (idx) idx[j] = mi;
        }
    };
    std::vector<std::thread> pool;
    const int64_t chunk = ((int64_t)a.q + a.nt - 1) / a.nt;
    for (int t = 0; t < a.nt; ++t) {
        const int64_t lo = t * chunk, hi = std::min<int64_t>(a.q, lo + chunk);
        if (lo < hi) pool.emplace_back(work, lo, hi);
    }
    for (auto &th : pool) th.join();
}

}  // namespace

int main(int argc, char **argv) {
    printf("----------------------------------\n");
    printf("   RMQ engine (MI355X, HIP)       \n");
    printf("----------------------------------\n");
    Args a;
    if (!parse(argc, argv, a)) {
        usage();
        return EXIT_FAILURE;
    }
    const bool idx_out = a.alg >= 100;
    const bool hrmq_alg = a.alg == ALG_CPU_HRMQ || a.alg == ALG_CPU_HRMQ_IDX;
    const bool cpu_alg = a.alg == ALG_CPU_BASE || a.alg == ALG_CPU_BASE_IDX || hrmq_alg;
    if (!cpu_alg && a.alg != ALG_GPU_BASE && a.alg != ALG_GPU_BASE_IDX && a.alg != ALG_GPU_RTX_BLOCKS &&
        a.alg != ALG_GPU_RTX_BLOCKS_IDX) {
        fprintf(stderr, "alg %d (%s) is an RT-core method outside this engine; use 5 / 105\n", a.alg,
                alg_name(a.alg));
        return 2;
    }
    printf("Params:\n   reps = %i\n   seed = %i\n   dev  = %i\n" C_GREEN "   n    = %i (~%f GB, float)\n" C_RESET
           "   bs   = %i\n" C_GREEN "   q    = %i (~%f GB, int2)\n" C_RESET "   lr   = %i\n   nt   = %i CPU threads\n"
           "   alg  = %i (%s)\n\n",
           a.reps, a.seed, a.dev, a.n, 4.0 * a.n / 1e9, a.bs, a.q, 8.0 * a.q / 1e9, a.lr, a.nt, a.alg,
           alg_name(a.alg));

    const bool need_gpu = !cpu_alg || a.check;
    if (need_gpu) {
        hip_ok(hipSetDevice(a.dev), "hipSetDevice");
        print_device(a.dev);
    }

    // 1) inputs
    double t0 = now_ms();
    printf(C_YELLOW "Generating n=%-10i values............", a.n);
    fflush(stdout);
    std::vector<float> hA;
    float *dA = nullptr;
    if (need_gpu) {
        hip_ok(hipMalloc(&dA, sizeof(float) * (size_t)a.n), "hipMalloc(A)");
        int rc = rmq_gen_array(dA, a.n, (uint64_t)a.seed, a.array_kind, nullptr);
        if (rc) die("rmq_gen_array", rc);
        hip_ok(hipDeviceSynchronize(), "generate");
    }
    if (cpu_alg || a.check) {
        hA.resize(a.n);
        if (dA) hip_ok(hipMemcpy(hA.data(), dA, sizeof(float) * (size_t)a.n, hipMemcpyDeviceToHost), "D2H A");
        else if (int rc = rmq_gen_array_host(hA.data(), a.n, (uint64_t)a.seed, a.array_kind)) die("rmq_gen_array_host", rc);
    }
    printf("done: %f secs\n", (now_ms() - t0) / 1000.0);
    t0 = now_ms();
    printf("Generating q=%-10i queries...........", a.q);
    fflush(stdout);
    std::vector<int32_t> hQ(2 * (size_t)a.q);
    if (a.qgen_ref) {
        if (int rc = rmq_gen_queries_ref(hQ.data(), a.q, a.n, a.lr, a.seed, a.nt)) die("rmq_gen_queries_ref", rc);
    } else if (int rc = rmq_gen_queries_host(hQ.data(), a.q, a.n, a.lr, (uint64_t)a.seed, a.nt)) {
        die("rmq_gen_queries_host", rc);
    }
    int32_t *dQ = nullptr;
    if (need_gpu) {
        hip_ok(hipMalloc(&dQ, sizeof(int32_t) * 2 * (size_t)a.q), "hipMalloc(Q)");
        hip_ok(hipMemcpy(dQ, hQ.data(), sizeof(int32_t) * 2 * (size_t)a.q, hipMemcpyHostToDevice), "H2D Q");
    }
    printf("done: %f secs\n" C_RESET, (now_ms() - t0) / 1000.0);

    const bool engine = a.alg == ALG_GPU_RTX_BLOCKS || a.alg == ALG_GPU_RTX_BLOCKS_IDX;
    rmq_query_opts_t opts{a.short_max, a.group_lanes, 0, a.long_lanes};
    if (engine && a.tune) {
        // timed passes over a prefix of the queries; answers do not depend on the pick
        // 2^24 queries: at 2^22 a timed pass takes ~20 us and the picks are noisy
        const int64_t sample = std::min<int64_t>(a.q, (int64_t)1 << 24);
        float *d_sv = nullptr;
        int32_t *d_si = nullptr;
        hip_ok(hipMalloc(&d_sv, sizeof(float) * (size_t)sample), "hipMalloc(tune)");
        hip_ok(hipMalloc(&d_si, sizeof(int32_t) * (size_t)sample), "hipMalloc(tune)");
        printf("Autotuning on %lld queries.................", (long long)sample);
        fflush(stdout);
        t0 = now_ms();
        rmq_tune_result_t pick{};
        const int32_t only_bs = a.bs;  // the block size given with --bs/--nb, the only candidate then
        if (int rc = rmq_tune(dA, a.n, dQ, sample, a.bs_given ? &only_bs : nullptr, a.bs_given ? 1 : 0,
                              idx_out ? nullptr : d_sv, idx_out ? d_si : nullptr,
                              a.tune_cache.empty() ? nullptr : a.tune_cache.c_str(), nullptr, &pick))
            die("rmq_tune", rc);
        (void)hipFree(d_sv);
        (void)hipFree(d_si);
        a.bs = pick.bs;
        a.nb = a.n / a.bs;
        a.edges = (pick.build_flags & RMQ_BUILD_EDGES) != 0;
        a.mid = (pick.build_flags & RMQ_BUILD_MID) != 0;
        opts = pick.opts;
        printf("done: %f secs [bs=%d edges=%d mid=%d short_max=%d group_lanes=%d long_lanes=%d; %s]\n",
               (now_ms() - t0) / 1000.0, a.bs, (int)a.edges, (int)a.mid, opts.short_max, opts.group_lanes,
               opts.long_lanes, pick.from_cache ? "cached" : "timed");
    }

    csv_prefix(a);

    // 2) computation
    std::vector<float> out_v(idx_out ? 0 : a.q);
    std::vector<int32_t> out_i(idx_out ? a.q : 0);
    double time_ms = 0, construction_ms = 0, out_bytes = 0, temp_bytes = 0;
    PowerSampler power;
    CpuEnergySampler cpu_power;
    if (cpu_alg) {
        // HRMQ (src/cpu_methods.h:80-159): the structure is built on one
        // thread and reported as construction; the values are the float
        // array's bits read as int32 (src/main.cu:117-125)
        rmq_hrmq_t *hr = nullptr;
        if (hrmq_alg) {
            printf("Creating MinMaxTree.......................");
            fflush(stdout);
            t0 = now_ms();
            if (int rc = rmq_hrmq_build(reinterpret_cast<const int32_t *>(hA.data()), a.n, &hr)) die("rmq_hrmq_build", rc);
            construction_ms = now_ms() - t0;
            int64_t bytes = 0;
            rmq_hrmq_bytes(hr, &bytes);
            out_bytes = (double)bytes;
            printf("done: %f ms (%f MB)\n", construction_ms, bytes / 1e6);
        }
        printf(C_CYAN "Computing RMQs%s (%-16s,nt=%2i,r=%-3i).." C_RESET, idx_out ? " index" : "", alg_name(a.alg), a.nt,
               a.reps);
        fflush(stdout);
        if (a.save_power) cpu_power.start(a.power_file, alg_name(a.alg));
        t0 = now_ms();
        for (int r = 0; r < a.reps; ++r) {
            if (hrmq_alg) {
                int rc = rmq_hrmq_query(hr, hQ.data(), a.q, idx_out ? out_i.data() : nullptr,
                                        idx_out ? nullptr : reinterpret_cast<int32_t *>(out_v.data()), a.nt);
                if (rc) die("rmq_hrmq_query", rc);
            } else {
                cpu_base(a, hA.data(), hQ.data(), idx_out ? nullptr : out_v.data(), idx_out ? out_i.data() : nullptr);
            }
        }
        time_ms = now_ms() - t0;
        rmq_hrmq_free(hr);
        if (a.save_power) {
            cpu_power.stop();
            if (cpu_power.energy_j() > 0)
                printf("energy efficiency: %.4e RMQ/J\n", (double)a.q * a.reps / cpu_power.energy_j());
        }
        print_rate("", a, time_ms);
    } else {
        float *d_val = nullptr;
        int32_t *d_idx = nullptr;
        if (idx_out) hip_ok(hipMalloc(&d_idx, sizeof(int32_t) * (size_t)a.q), "hipMalloc(out)");
        else hip_ok(hipMalloc(&d_val, sizeof(float) * (size_t)a.q), "hipMalloc(out)");
        rmq_handle_t *h = nullptr;
        if (engine) {
            printf("Building block index (bs=%i%s%s)........", a.bs, a.edges ? ", edge tables" : "",
                   a.edges && a.mid ? ", coarse rows" : "");
            fflush(stdout);
            t0 = now_ms();
            const uint32_t flags = a.edges ? (RMQ_BUILD_EDGES | (a.mid ? RMQ_BUILD_MID : 0u)) : 0u;
            if (int rc = rmq_build(dA, a.n, a.bs, flags, nullptr, &h)) die("rmq_build", rc);
            hip_ok(hipDeviceSynchronize(), "build");
            construction_ms = now_ms() - t0;
            int64_t nb = 0, bytes = 0;
            int32_t levels = 0;
            rmq_handle_info(h, nullptr, nullptr, &nb, &levels, &bytes);
            out_bytes = (double)bytes;
            temp_bytes = 256.0 + 4.0 * a.q;
            printf("done: %f ms [%lld blocks, %d levels, %f MB]\n", construction_ms, (long long)nb, levels,
                   bytes / 1e6);
        }
        printf(C_CYAN "Computing RMQs%s (%-16s,r=%-3i)..." C_RESET, idx_out ? " index" : "", alg_name(a.alg), a.reps);
        fflush(stdout);
        if (a.save_power) power.start(a.dev, a.power_file, alg_name(a.alg));
        t0 = now_ms();
        for (int r = 0; r < a.reps; ++r) {  // src/rtx.h:151-154: launch, then synchronise, per rep
            int rc = engine ? rmq_query_ex(h, dQ, a.q, d_val, d_idx, &opts, nullptr)
                            : rmq_gpu_base(dA, a.n, dQ, a.q, d_val, d_idx, nullptr);
            if (rc) die(engine ? "rmq_query" : "rmq_gpu_base", rc);
            hip_ok(hipDeviceSynchronize(), "query");
        }
        time_ms = now_ms() - t0;
        if (a.save_power) {
            power.stop();
            // queries answered per joule over the query loop, q * reps / energy
            // (scripts/plot_rmq_per_joule.py:61-76)
            if (power.energy_j() > 0)
                printf("energy efficiency: %.4e RMQ/J\n", (double)a.q * a.reps / power.energy_j());
        }
        print_rate("", a, time_ms);
        if (idx_out) hip_ok(hipMemcpy(out_i.data(), d_idx, sizeof(int32_t) * (size_t)a.q, hipMemcpyDeviceToHost), "D2H");
        else hip_ok(hipMemcpy(out_v.data(), d_val, sizeof(float) * (size_t)a.q, hipMemcpyDeviceToHost), "D2H");
        rmq_free(h);
        (void)hipFree(d_val);
        (void)hipFree(d_idx);
    }
    csv_suffix(a, time_ms, construction_ms, out_bytes, temp_bytes);

    // 3) --check against the GPU brute force (src/main.cu:140-167), bit for bit
    int status = EXIT_SUCCESS;
    if (a.check) {
        printf("\nCHECKING RESULT:\n");
        std::vector<float> ev(a.q);
        std::vector<int32_t> ei(a.q);
        float *d_ev = nullptr;
        int32_t *d_ei = nullptr;
        hip_ok(hipMalloc(&d_ev, sizeof(float) * (size_t)a.q), "hipMalloc");
        hip_ok(hipMalloc(&d_ei, sizeof(int32_t) * (size_t)a.q), "hipMalloc");
        if (int rc = rmq_gpu_base(dA, a.n, dQ, a.q, d_ev, d_ei, nullptr)) die("rmq_gpu_base", rc);
        hip_ok(hipMemcpy(ev.data(), d_ev, sizeof(float) * (size_t)a.q, hipMemcpyDeviceToHost), "D2H");
        hip_ok(hipMemcpy(ei.data(), d_ei, sizeof(int32_t) * (size_t)a.q, hipMemcpyDeviceToHost), "D2H");
        printf(C_YELLOW "Checking result..........................." C_RESET);
        int64_t bad = 0;
        // HRMQ's index is the rightmost minimum (alg 101): it must hold the
        // same value as the leftmost one GPU BASE returns and lie in [l, r]
        for (int64_t j = 0; j < a.q; ++j) {
            const int32_t l = hQ[2 * j], r = hQ[2 * j + 1];
            const bool ok = !idx_out ? memcmp(&out_v[j], &ev[j], 4) == 0
                            : a.alg == ALG_CPU_HRMQ_IDX
                                ? out_i[j] >= l && out_i[j] <= r && memcmp(&hA[out_i[j]], &ev[j], 4) == 0
                                : out_i[j] == ei[j];
            if (!ok && bad++ < 10) {
                if (idx_out) printf("\nError on %lld-th query [%d,%d]: got %d, expected %d", (long long)j, hQ[2 * j],
                                    hQ[2 * j + 1], out_i[j], ei[j]);
                else printf("\nError on %lld-th query [%d,%d]: got %.9g, expected %.9g", (long long)j, hQ[2 * j],
                            hQ[2 * j + 1], out_v[j], ev[j]);
            }
        }
        printf(C_YELLOW "%s" C_RESET " (%lld mismatches)\n", bad ? "failed" : "pass", (long long)bad);
        if (bad) status = 3;
        (void)hipFree(d_ev);
        (void)hipFree(d_ei);
    }
    (void)hipFree(dA);
    (void)hipFree(dQ);
    printf("Benchmark Finished\n");
    return status;
}
