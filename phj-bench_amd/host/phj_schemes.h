// phj_schemes.h — drop-in scheme entry points backed by the MI355X library.
//
// Signatures and result structs match src/CHJ.h:12-31, src/PHJ.h:13-31 and
// src/PHJBep.h:13-98. Inputs are copied to the device, joined there, and the
// output comes back as 10,000-row OutputBlocks (src/JoinOutput.h:18-77) in one
// OutputWorker. Errors are thrown as in the reference: std::invalid_argument
// for bad radix configurations or schemas, std::bad_alloc when memory runs
// out, std::runtime_error for device failures.
#pragma once

#include <cstddef>

#include "phj_types.h"

namespace phj {

struct ChjResult {
    PhaseTiming build;
    PhaseTiming probe;
    double e2e_wall_ms = 0.0;
    size_t peak_mem_bytes = 0;
    JoinOutput output;
};

struct PhjResult {
    PhaseTiming build_shuffle;
    PhaseTiming build;
    PhaseTiming probe_shuffle;
    PhaseTiming probe;
    double e2e_wall_ms = 0.0;
    size_t peak_mem_bytes = 0;
    JoinOutput output;
};

struct PhjBepResult {
    PhaseTiming build_shuffle;
    PhaseTiming build;
    PhaseTiming probe_shuffle;
    PhaseTiming probe;
    PhaseTiming eviction_overhead;
    double e2e_wall_ms = 0.0;
    size_t peak_mem_bytes = 0;
    JoinOutput output;
    size_t bep_budget_mib = 0;
    size_t bep_evictions = 0;
    size_t bep_refinements = 0;
    size_t bep_peak_bytes = 0;
    size_t bep_build_skip_retries = 0;
};

ChjResult runCHJ(const BlockStream &build, const BlockStream &probe, size_t threads);
PhjResult runPHJ(const BlockStream &build, const BlockStream &probe, const RadixConfig &cfg, size_t threads);
PhjBepResult runPhjBep(const BlockStream &build, const BlockStream &probe, const RadixConfig &cfg, size_t threads,
                       size_t bep_budget_mib);

} // namespace phj
