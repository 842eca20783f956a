// Basis context: the device-resident counterpart of the reference's
// LibcintContext (cpp/include/ao_eval.hpp:20-100, cpp/src/ao_eval.cpp:78-158).
#pragma once

#include "common.hpp"

#include <map>
#include <memory>
#include <mutex>

namespace mpisdf {

// AO chunk for the evaluation kernel: shells [s0, s1) own AOs [a0, a1).
struct AoChunk {
    int s0, s1, a0, a1;
};

struct Context {
    // Private copies of the tables (the reference context owns copies too,
    // python/test/test_ao_eval.py:51-70).
    std::vector<int32_t> atm, bas;
    std::vector<double> env;
    bool cart = false;
    int64_t nao = 0;
    int nshells = 0;
    int lmax = 0;
    int max_shell_width = 0;  // nctr * ncomp of the widest shell
    std::vector<DevShell> shells;
    std::vector<double> tabs;       // per-l transforms, ncomp x ncart row-major
    std::vector<int> tab_offset;    // offset of l's table in tabs

    DeviceBuffer<DevShell> d_shells;
    DeviceBuffer<double> d_env;
    DeviceBuffer<double> d_tabs;

    // The randomized selectors' sketch storage, kept between selections and
    // grown to the largest one asked for: a sketch of ~200 GB freed and
    // allocated again would wait seconds for the driver to clear the freed
    // memory first.  Selections on one context take sketch_mutex while they
    // use it; release_sketch_store() gives the memory back.
    std::mutex sketch_mutex;
    DeviceBuffer<uint8_t> sketch_store;
    void* reserve_sketch(size_t bytes);
    void release_sketch_store();

    // Chunk tables per chunk width (built lazily, one per kernel variant).
    std::mutex chunk_mutex;
    std::map<int, std::pair<std::vector<AoChunk>, std::unique_ptr<DeviceBuffer<AoChunk>>>> chunks;

    const std::pair<std::vector<AoChunk>, std::unique_ptr<DeviceBuffer<AoChunk>>>& chunk_table(int width);
};

// Validates ATM/BAS/ENV exactly as cpp/src/ao_eval.cpp:54-120 does and
// builds shells, transforms and culling radii; uploads to the current device.
std::unique_ptr<Context> make_context(const int32_t* atm, int64_t natm, const int32_t* bas,
                                      int64_t nbas, const double* env, int64_t nenv, bool cart);

// Real-solid-harmonic transform (ncomp x ncart) for one angular momentum.
std::vector<double> transform_table(int l, bool cart);

// Analytic overlap S[mu * nao + nu] of the context's AOs (overlap.cpp).
void overlap_matrix(const Context& ctx, double* S);

}  // namespace mpisdf
