// phj_bench — the phj-bench command line, run on the MI355X library.
//
// Options, defaults, validation rules, exit codes and the [info]/[check] lines
// follow the reference driver (src/CLI.cpp:17-355, src/Main.cpp:15-164); the
// stdout summary and the per-rep CSV follow its report (src/Report.cpp:14-483),
// column for column, so runs of either program land in one spreadsheet.
//
// What differs from the reference driver:
//   - Both relations are generated in device memory by the library's
//     reference-stream generator (PHJ_DIST_REFERENCE: the default build and
//     probe streams of src/DataGen.cpp:84-139, every probe row matching) and
//     the joins run on them there; generation stays outside the timed window
//     as in src/Main.cpp:27-36.
//   - --check compares each scheme's first-rep output checksum (per-row digest
//     of key | build cols | probe cols, summed; src/Reference.cpp:22-43 defines
//     the row bytes) with the checksum the generator implies, instead of
//     building a std::unordered_map reference on the host.
//   - --threads is recorded in the report; the device does not use it.
//   - --device N selects the GPU (default 0); the reference has no devices.
#include <algorithm>
#include <bit>
#include <charconv>
#include <chrono>
#include <cstdint>
#include <cstdio>
#include <fstream>
#include <iomanip>
#include <iostream>
#include <optional>
#include <sstream>
#include <string>
#include <string_view>
#include <vector>

#include "phj_gpu.h"

namespace {

const char *kUsage = R"(phj_bench: CHJ vs PHJ vs PHJ-PURE hash joins on one MI355X

Usage:
  phj_bench [options]

Options:
  --scheme {chj|phj|phj-pure|all} Scheme(s) to run (default: chj and phj;
                                 all adds phj-pure).
  --build-rows N                 Build-side rows (default: 1000000).
  --probe-rows N                 Probe-side rows (default: 1000000).
  --build-payload-schema LIST    Build payload types, comma separated (u32,u64,...).
  --probe-payload-schema LIST    Probe payload types, comma separated.
  --threads N                    Recorded in the report (default: 1); the device
                                 does not use it.
  --partitions N                 Total partitions, a power of 2 (default: 256).
  --passes K                     Radix passes (default: 1); the partition bits
                                 are split evenly, earlier passes taking the rest.
  --partition-bits-per-pass LIST Bits of each pass, e.g. 8,8 (overrides
                                 --partitions and --passes).
  --bep-budget-mib N             PHJ probe buffer budget in MiB (default: 512).
  --reps N                       Repetitions per scheme (default: 1).
  --csv PATH                     Append one CSV row per rep to PATH (header
                                 written when the file is new).
  --check                        Check each scheme's output against the join
                                 the generator implies.
  --seed N                       Generator seed (default: 42).
  --device N                     GPU ordinal (default: 0).
  -h, --help                     Print this message and exit.

Types: u8, u16, u32, u64, u128.
)";

// src/CLI.h:15-22
enum class Scheme { Default = 0, CHJ = 1, PHJ = 2, PhjPure = 3, All = 4 };

struct Options {
    Scheme scheme = Scheme::Default;
    uint64_t build_rows = 1000000;
    uint64_t probe_rows = 1000000;
    std::vector<uint8_t> build_types;
    std::vector<uint8_t> probe_types;
    uint64_t threads = 1;
    std::vector<uint8_t> pass_bits{8};
    uint64_t reps = 1;
    std::string csv_path;
    bool check = false;
    uint64_t seed = 42;
    uint64_t bep_budget_mib = 512;
    int device = 0;

    uint32_t total_bits() const {
        uint32_t t = 0;
        for (uint8_t b : pass_bits)
            t += b;
        return t;
    }
    uint64_t partitions() const { return total_bits() >= 64 ? 0 : uint64_t{1} << total_bits(); }
};

const char *const kTypeNames[] = {"u8", "u16", "u32", "u64", "u128"};

bool parse_u64(std::string_view s, uint64_t &out) {
    uint64_t v = 0;
    const auto r = std::from_chars(s.data(), s.data() + s.size(), v);
    if (r.ec != std::errc{} || r.ptr != s.data() + s.size() || s.empty())
        return false;
    out = v;
    return true;
}

// Comma-separated items, each parsed by f. As src/CLI.cpp:87-130 does, a
// trailing comma ends the list and an empty string is an empty list (which
// the bits list then rejects).
template <class F> bool parse_list(std::string_view s, F f) {
    size_t i = 0;
    while (i < s.size()) {
        size_t j = s.find(',', i);
        if (j == std::string_view::npos)
            j = s.size();
        if (!f(s.substr(i, j - i)))
            return false;
        i = j + 1;
    }
    return true;
}

bool parse_schema(std::string_view s, std::vector<uint8_t> &out) {
    out.clear();
    return parse_list(s, [&](std::string_view item) {
        for (uint8_t t = 0; t < 5; ++t)
            if (item == kTypeNames[t]) {
                out.push_back(t);
                return true;
            }
        return false;
    });
}

bool parse_bits(std::string_view s, std::vector<uint8_t> &out) {
    out.clear();
    return parse_list(s, [&](std::string_view item) {
        uint64_t v = 0;
        if (!parse_u64(item, v) || v == 0 || v > 16)
            return false;
        out.push_back(static_cast<uint8_t>(v));
        return true;
    }) && !out.empty();
}

// Returns the options, or nullopt with rc set: 0 after --help, 2 on errors.
std::optional<Options> parse_cli(int argc, char **argv, int &rc) {
    Options o;
    uint64_t parts = 0, passes = 0;
    std::vector<uint8_t> bits;
    auto fail = [&](const std::string &msg) {
        std::cerr << msg << "\n";
        rc = 2;
        return std::nullopt;
    };
    for (int i = 1; i < argc; ++i) {
        std::string_view arg = argv[i], key = arg, val;
        if (const size_t eq = arg.find('='); eq != std::string_view::npos) {
            key = arg.substr(0, eq);
            val = arg.substr(eq + 1);
        } else if (arg != "--check" && arg != "--help" && arg != "-h") {
            if (i + 1 >= argc)
                return fail("missing value for " + std::string(key));
            val = argv[++i];
        }
        if (key == "--help" || key == "-h") {
            std::cout << kUsage;
            rc = 0;
            return std::nullopt;
        } else if (key == "--scheme") {
            if (val == "chj")
                o.scheme = Scheme::CHJ;
            else if (val == "phj")
                o.scheme = Scheme::PHJ;
            else if (val == "phj-pure")
                o.scheme = Scheme::PhjPure;
            else if (val == "all")
                o.scheme = Scheme::All;
            else
                return fail("--scheme must be chj|phj|phj-pure|all");
        } else if (key == "--build-rows") {
            if (!parse_u64(val, o.build_rows))
                return fail("invalid --build-rows");
        } else if (key == "--probe-rows") {
            if (!parse_u64(val, o.probe_rows))
                return fail("invalid --probe-rows");
        } else if (key == "--build-payload-schema") {
            if (!parse_schema(val, o.build_types))
                return fail("invalid --build-payload-schema");
        } else if (key == "--probe-payload-schema") {
            if (!parse_schema(val, o.probe_types))
                return fail("invalid --probe-payload-schema");
        } else if (key == "--threads") {
            if (!parse_u64(val, o.threads) || o.threads == 0)
                return fail("invalid --threads");
        } else if (key == "--partitions") {
            if (!parse_u64(val, parts) || parts == 0 || (parts & (parts - 1)) != 0)
                return fail("--partitions must be a positive power of 2");
        } else if (key == "--passes") {
            if (!parse_u64(val, passes) || passes == 0)
                return fail("invalid --passes");
        } else if (key == "--partition-bits-per-pass") {
            if (!parse_bits(val, bits))
                return fail("invalid --partition-bits-per-pass");
        } else if (key == "--reps") {
            if (!parse_u64(val, o.reps) || o.reps == 0)
                return fail("invalid --reps");
        } else if (key == "--csv") {
            o.csv_path = std::string(val);
        } else if (key == "--check") {
            o.check = true;
        } else if (key == "--seed") {
            if (!parse_u64(val, o.seed))
                return fail("invalid --seed");
        } else if (key == "--bep-budget-mib") {
            if (!parse_u64(val, o.bep_budget_mib))
                return fail("invalid --bep-budget-mib");
        } else if (key == "--device") {
            uint64_t d = 0;
            if (!parse_u64(val, d) || d > 255)
                return fail("invalid --device");
            o.device = static_cast<int>(d);
        } else {
            return fail("unknown option: " + std::string(key));
        }
    }
    // Radix configuration (src/CLI.cpp:323-352): an explicit bits list wins;
    // otherwise log2(partitions) bits split over the passes, the first
    // (total % passes) passes taking one bit more.
    if (!bits.empty()) {
        o.pass_bits = bits;
    } else {
        const uint32_t total = static_cast<uint32_t>(std::bit_width(parts == 0 ? uint64_t{256} : parts)) - 1u;
        const uint64_t k = passes == 0 ? 1 : passes;
        if (k > total)
            return fail("passes cannot exceed log2(partitions)");
        o.pass_bits.assign(k, static_cast<uint8_t>(total / k));
        for (uint64_t p = 0; p < total % k; ++p)
            ++o.pass_bits[p];
    }
    if (o.total_bits() == 0 || o.total_bits() > 32)
        return fail("total partition bits must be in [1, 32]");
    if (o.pass_bits.size() > PHJ_MAX_PASSES)
        return fail("at most " + std::to_string(PHJ_MAX_PASSES) + " passes are supported");
    if (o.build_types.size() > PHJ_MAX_COLS || o.probe_types.size() > PHJ_MAX_COLS)
        return fail("at most " + std::to_string(PHJ_MAX_COLS) + " payload columns per side are supported");
    return o;
}

std::string schema_string(const std::vector<uint8_t> &types) {
    if (types.empty())
        return "<none>";
    std::string s;
    for (size_t i = 0; i < types.size(); ++i)
        s += (i ? "," : "") + std::string(kTypeNames[types[i]]);
    return s;
}

std::string bits_string(const std::vector<uint8_t> &bits) {
    std::string s;
    for (size_t i = 0; i < bits.size(); ++i)
        s += (i ? "+" : "") + std::to_string(bits[i]);
    return s;
}

double mib(uint64_t bytes) { return static_cast<double>(bytes) / (1024.0 * 1024.0); }

// One rep of one scheme: the result's timings and counters (its device
// output is released as soon as the rep is logged and, when asked, checked).
struct Rep {
    phj_phase_timing build_shuffle{}, build{}, probe_shuffle{}, probe{}, eviction{};
    double e2e_ms = 0;
    uint64_t peak_bytes = 0, out_rows = 0;
    uint64_t bep_budget_mib = 0, bep_evictions = 0, bep_refinements = 0, bep_peak_bytes = 0, bep_skip_retries = 0;
};

struct Stats {
    double med = 0, min = 0, max = 0;
    bool present = false;
};

Stats stats(std::vector<double> v) {
    Stats s;
    if (v.empty())
        return s;
    std::sort(v.begin(), v.end());
    s.present = true;
    s.min = v.front();
    s.max = v.back();
    const size_t m = v.size() / 2;
    s.med = v.size() % 2 ? v[m] : 0.5 * (v[m - 1] + v[m]);
    return s;
}

template <class F> Stats over(const std::vector<Rep> &reps, F f) {
    std::vector<double> v;
    for (const Rep &r : reps)
        v.push_back(f(r));
    return stats(std::move(v));
}

std::string fixed(double x, int prec) {
    std::ostringstream s;
    s << std::fixed << std::setprecision(prec) << x;
    return s.str();
}

// The per-scheme table of src/Report.cpp:79-172: a 16-wide phase column and
// six 14-wide cells (wall ms and ns/row, each median/min/max).
void print_table(std::ostream &os, const char *name, const std::vector<Rep> &reps, bool shuffles, bool bep) {
    constexpr int W = 14, P = 16;
    const std::string rule(P + 6 * W, '-');
    auto phase = [&](const std::string &s) { os << std::left << std::setw(P) << s << std::right; };
    auto cell = [&](const std::string &s) { os << std::setw(W) << s; };
    auto three = [&](const Stats &s, int prec) {
        cell(fixed(s.med, prec));
        cell(fixed(s.min, prec));
        cell(fixed(s.max, prec));
    };
    auto blanks = [&](int n) {
        for (int i = 0; i < n; ++i)
            cell("");
    };
    os << "\nScheme: " << name << "\n" << rule << "\n";
    phase("phase");
    for (const char *h : {"wall_ms_med", "wall_ms_min", "wall_ms_max", "ns/row_med", "ns/row_min", "ns/row_max"})
        cell(h);
    os << "\n" << rule << "\n";
    auto row = [&](const char *label, phj_phase_timing Rep::*m) {
        phase(label);
        three(over(reps, [&](const Rep &r) { return (r.*m).wall_ms; }), 3);
        three(over(reps, [&](const Rep &r) { return (r.*m).ns_per_row; }), 2);
        os << "\n";
    };
    if (shuffles)
        row("build-shuffle", &Rep::build_shuffle);
    row("build", &Rep::build);
    if (shuffles)
        row("probe-shuffle", &Rep::probe_shuffle);
    row("probe", &Rep::probe);
    if (bep)
        row("eviction-overhead", &Rep::eviction);
    phase("e2e_wall_ms");
    three(over(reps, [](const Rep &r) { return r.e2e_ms; }), 3);
    blanks(3);
    os << "\n";
    phase("peak_mem_mib");
    three(over(reps, [](const Rep &r) { return mib(r.peak_bytes); }), 1);
    blanks(3);
    os << "\n" << rule << "\n";
    if (bep) {
        phase("bep_budget_mib");
        cell(fixed(static_cast<double>(reps.front().bep_budget_mib), 0));
        blanks(5);
        os << "\n";
        auto metric = [&](const char *label, double (*f)(const Rep &), int prec) {
            phase(label);
            three(over(reps, f), prec);
            blanks(3);
            os << "\n";
        };
        metric("bep_evictions", [](const Rep &r) { return static_cast<double>(r.bep_evictions); }, 0);
        metric("bep_refinements", [](const Rep &r) { return static_cast<double>(r.bep_refinements); }, 0);
        metric("bep_peak_mib", [](const Rep &r) { return mib(r.bep_peak_bytes); }, 3);
        metric("bep_skip_retries", [](const Rep &r) { return static_cast<double>(r.bep_skip_retries); }, 0);
        os << rule << "\n";
    }
}

// src/Report.cpp:312-361
void print_summary(std::ostream &os, const Options &o, const std::vector<Rep> &chj, const std::vector<Rep> &pure,
                   const std::vector<Rep> &bep) {
    os << "== Run configuration ==\n";
    auto line = [&](const char *k, const std::string &v) { os << "  " << std::left << std::setw(24) << k << v << "\n"; };
    line("threads:", std::to_string(o.threads));
    line("build_rows:", std::to_string(o.build_rows));
    line("probe_rows:", std::to_string(o.probe_rows));
    line("build_payload_schema:", schema_string(o.build_types));
    line("probe_payload_schema:", schema_string(o.probe_types));
    line("partitions (PHJ):", std::to_string(o.partitions()) + " (bits per pass: " + bits_string(o.pass_bits) + ")");
    line("bep_budget_mib:", std::to_string(o.bep_budget_mib));
    line("reps:", std::to_string(o.reps));
    line("distribution:", "uniform");
    line("match_rate:", "1.000");
    line("seed:", std::to_string(o.seed));
    os << std::right;
    if (!chj.empty())
        print_table(os, "CHJ", chj, false, false);
    if (!pure.empty())
        print_table(os, "PHJ-PURE", pure, true, false);
    if (!bep.empty())
        print_table(os, "PHJ", bep, true, true);
}

// The CSV of src/Report.cpp:364-483: appended, header only for a new file;
// columns a scheme does not have are left empty.
const char *kCsvHeader =
    "scheme,rep,threads,build_rows,probe_rows,build_payload_schema,probe_payload_schema,partitions,distribution,"
    "match_rate,build_wall_ms,build_ns_per_row,probe_wall_ms,probe_ns_per_row,build_shuffle_wall_ms,"
    "build_shuffle_ns_per_row,probe_shuffle_wall_ms,probe_shuffle_ns_per_row,eviction_overhead_wall_ms,"
    "eviction_overhead_ns_per_row,e2e_wall_ms,peak_mem_mib,bep_budget_mib,bep_evictions,bep_refinements,"
    "bep_peak_mib,bep_build_skip_retries";

void write_csv(const Options &o, const std::vector<Rep> &chj, const std::vector<Rep> &pure,
               const std::vector<Rep> &bep) {
    const bool fresh = !std::ifstream(o.csv_path).good();
    std::ofstream f(o.csv_path, std::ios::out | std::ios::app);
    if (!f.good())
        return;
    if (fresh)
        f << kCsvHeader << "\n";
    f << std::fixed << std::setprecision(6);
    auto emit = [&](const char *scheme, const std::vector<Rep> &reps, bool shuffles, bool is_bep) {
        for (size_t i = 0; i < reps.size(); ++i) {
            const Rep &r = reps[i];
            f << scheme << ',' << i << ',' << o.threads << ',' << o.build_rows << ',' << o.probe_rows << ",\""
              << schema_string(o.build_types) << "\",\"" << schema_string(o.probe_types) << "\"," << o.partitions()
              << ",uniform,1.0,";
            auto val = [&](double v) { f << v << ','; };
            auto blank = [&](int n) {
                for (int k = 0; k < n; ++k)
                    f << ',';
            };
            val(r.build.wall_ms);
            val(r.build.ns_per_row);
            val(r.probe.wall_ms);
            val(r.probe.ns_per_row);
            if (shuffles) {
                val(r.build_shuffle.wall_ms);
                val(r.build_shuffle.ns_per_row);
                val(r.probe_shuffle.wall_ms);
                val(r.probe_shuffle.ns_per_row);
            } else {
                blank(4);
            }
            if (is_bep) {
                val(r.eviction.wall_ms);
                val(r.eviction.ns_per_row);
            } else {
                blank(2);
            }
            val(r.e2e_ms);
            val(mib(r.peak_bytes));
            if (is_bep)
                f << r.bep_budget_mib << ',' << r.bep_evictions << ',' << r.bep_refinements << ','
                  << mib(r.bep_peak_bytes) << ',' << r.bep_skip_retries;
            else
                blank(4);
            f << '\n';
        }
    };
    emit("CHJ", chj, false, false);
    emit("PHJ-PURE", pure, true, false);
    emit("PHJ", bep, true, true);
}

double ms_since(std::chrono::steady_clock::time_point t0) {
    return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
}

struct Failure {
    std::string msg;
};

void ok_or_throw(phj_ctx *ctx, phj_status rc, const char *what) {
    if (rc != PHJ_OK)
        throw Failure{std::string(what) + ": " + (ctx ? phj_last_error(ctx) : "no context") + " (status " +
                      std::to_string(static_cast<int>(rc)) + ")"};
}

int run(const Options &o) {
    phj_ctx *ctx = nullptr;
    if (phj_ctx_create(o.device, &ctx) != PHJ_OK) {
        std::cerr << "[error] no usable HIP device " << o.device << " for the hash-join library\n";
        return 1;
    }
    struct Guard {
        phj_ctx *c;
        ~Guard() { phj_ctx_destroy(c); }
    } guard{ctx};

    phj_gen_spec spec{};
    spec.seed = o.seed;
    spec.build_rows = o.build_rows;
    spec.probe_rows = o.probe_rows;
    spec.distribution = PHJ_DIST_REFERENCE;
    spec.match_ppm = 1000000;
    spec.build_n_cols = static_cast<uint32_t>(o.build_types.size());
    spec.probe_n_cols = static_cast<uint32_t>(o.probe_types.size());
    std::copy(o.build_types.begin(), o.build_types.end(), spec.build_types);
    std::copy(o.probe_types.begin(), o.probe_types.end(), spec.probe_types);

    phj_relation build{}, probe{};
    auto t0 = std::chrono::steady_clock::now();
    ok_or_throw(ctx, phj_gen_relations(ctx, &spec, &build, &probe), "data generation");
    ok_or_throw(ctx, phj_synchronize(ctx), "data generation");
    const uint64_t blk = PHJ_PIPELINE_BLOCK_ROWS;
    std::cerr << "[info] data generation: " << ms_since(t0) << " ms  (build_rows=" << o.build_rows
              << ", probe_rows=" << o.probe_rows << ", threads=" << o.threads
              << ", blocks(build)=" << (o.build_rows + blk - 1) / blk
              << ", blocks(probe)=" << (o.probe_rows + blk - 1) / blk << ")\n";

    phj_checksum want{};
    if (o.check) {
        t0 = std::chrono::steady_clock::now();
        ok_or_throw(ctx, phj_gen_expected_checksum(ctx, &spec, &want), "expected checksum");
        std::cerr << "[info] reference build: " << ms_since(t0) << " ms  (output_rows=" << want.rows << ")\n";
    }

    const bool run_chj = o.scheme == Scheme::CHJ || o.scheme == Scheme::All || o.scheme == Scheme::Default;
    const bool run_pure = o.scheme == Scheme::PhjPure || o.scheme == Scheme::All;
    const bool run_bep = o.scheme == Scheme::PHJ || o.scheme == Scheme::All || o.scheme == Scheme::Default;

    std::vector<Rep> chj, pure, bep;
    std::vector<std::string> check_lines;
    int rc = 0;
    auto run_scheme = [&](phj_scheme scheme, const char *name, std::vector<Rep> &reps) {
        phj_join_desc d{};
        d.scheme = scheme;
        d.threads = static_cast<uint32_t>(std::min<uint64_t>(o.threads, UINT32_MAX));
        d.build = build;
        d.probe = probe;
        d.n_passes = static_cast<uint32_t>(o.pass_bits.size());
        std::copy(o.pass_bits.begin(), o.pass_bits.end(), d.pass_bits);
        d.bep_budget_bytes = o.bep_budget_mib << 20;
        for (uint64_t i = 0; i < o.reps; ++i) {
            phj_join_result r{};
            ok_or_throw(ctx, phj_join(ctx, &d, &r), name);
            Rep x;
            x.build_shuffle = r.build_shuffle;
            x.build = r.build;
            x.probe_shuffle = r.probe_shuffle;
            x.probe = r.probe;
            x.eviction = r.eviction_overhead;
            x.e2e_ms = r.e2e_wall_ms;
            x.peak_bytes = r.peak_mem_bytes;
            x.out_rows = r.out_rows;
            x.bep_budget_mib = r.bep_budget_mib;
            x.bep_evictions = r.bep_evictions;
            x.bep_refinements = r.bep_refinements;
            x.bep_peak_bytes = r.bep_peak_bytes;
            x.bep_skip_retries = r.bep_build_skip_retries;
            std::cerr << "[info] " << name << " rep " << i << ":";
            if (scheme == PHJ_SCHEME_CHJ)
                std::cerr << " build " << x.build.wall_ms << " ms, probe " << x.probe.wall_ms << " ms";
            else if (scheme == PHJ_SCHEME_PHJ_PURE)
                std::cerr << " shuffle " << x.build_shuffle.wall_ms << "/" << x.probe_shuffle.wall_ms << " ms, build "
                          << x.build.wall_ms << " ms, probe " << x.probe.wall_ms << " ms";
            else
                std::cerr << " shuffle(build) " << x.build_shuffle.wall_ms << " ms, build " << x.build.wall_ms
                          << " ms, shuffle(probe) " << x.probe_shuffle.wall_ms << " ms, probe " << x.probe.wall_ms
                          << " ms, evict " << x.eviction.wall_ms << " ms";
            std::cerr << ", e2e " << x.e2e_ms << " ms, out_rows " << x.out_rows;
            if (scheme == PHJ_SCHEME_PHJ)
                std::cerr << ", budget " << x.bep_budget_mib << " MiB, evictions " << x.bep_evictions
                          << ", refinements " << x.bep_refinements << ", bep_peak_mib " << mib(x.bep_peak_bytes)
                          << " MiB, skip-retries " << x.bep_skip_retries;
            std::cerr << "\n";
            if (o.check && i == 0) {
                phj_checksum got{};
                const phj_status st = phj_result_checksum(ctx, &r, &got);
                if (st != PHJ_OK) {
                    phj_result_release(ctx, &r);
                    ok_or_throw(ctx, st, name);
                }
                if (got.rows == want.rows && got.lane0 == want.lane0 && got.lane1 == want.lane1) {
                    check_lines.push_back(std::string("[check] ") + name + " OK (" + std::to_string(want.rows) +
                                          " rows)");
                } else {
                    std::ostringstream m;
                    m << "[check] " << name << " FAILED: " << got.rows << " rows, checksum " << std::hex << got.lane0
                      << ":" << got.lane1 << " vs " << std::dec << want.rows << " rows, checksum " << std::hex
                      << want.lane0 << ":" << want.lane1;
                    check_lines.push_back(m.str());
                    rc = 1;
                }
            }
            phj_result_release(ctx, &r);
            reps.push_back(x);
        }
    };
    try {
        if (run_chj)
            run_scheme(PHJ_SCHEME_CHJ, "CHJ", chj);
        if (run_pure)
            run_scheme(PHJ_SCHEME_PHJ_PURE, "PHJ-PURE", pure);
        if (run_bep)
            run_scheme(PHJ_SCHEME_PHJ, "PHJ", bep);
    } catch (...) {
        phj_relation_release(ctx, &build);
        phj_relation_release(ctx, &probe);
        throw;
    }
    phj_relation_release(ctx, &build);
    phj_relation_release(ctx, &probe);
    for (const std::string &l : check_lines)
        std::cerr << l << "\n";
    if (rc != 0)
        return rc;
    print_summary(std::cout, o, chj, pure, bep);
    if (!o.csv_path.empty())
        write_csv(o, chj, pure, bep);
    return 0;
}

} // namespace

int main(int argc, char **argv) {
    int rc = 0;
    const std::optional<Options> o = parse_cli(argc, argv, rc);
    if (!o)
        return rc;
    try {
        return run(*o);
    } catch (const Failure &f) {
        std::cerr << "[error] " << f.msg << "\n";
        return 1;
    }
}
