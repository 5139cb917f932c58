// Test infrastructure only: exposes the reference's own StableUnitRandom
// (src/native/legacy/source_file_realtime_v1_async/LearningRule/src/Rule/CerebullarLearningRule.cpp:10-24)
// behind a C symbol. The function body is the reference's text, cut out of that file at build
// time into _ref/stable_unit_random.inc by the Makefile's `ref` target; nothing of it is kept
// in this repository. Used only to write and check tests/golden/stable_unit_random.npz.
#include <cstdint>

namespace ref_cerebellar {
#include "stable_unit_random.inc"
}

extern "C" float ref_stable_unit_random(unsigned int seed, int rule_id, int synapse_id, int spike_time) {
    return ref_cerebellar::StableUnitRandom(seed, rule_id, synapse_id, spike_time);
}
