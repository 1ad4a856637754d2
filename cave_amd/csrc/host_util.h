// host_util.h -- host-side helpers shared by the C-ABI implementations.
#pragma once

#include <cstdint>
#include <string>
#include <vector>

#include "common.h"
#include "device_util.h"

namespace cave {

std::vector<uint64_t> enumerate_valid_bitmasks(uint32_t n, uint32_t k);
uint32_t neighbor_count(uint32_t flags);
uint64_t rule_string_to_bitmask(const std::string &rule);
std::string bitmask_to_rule_string(uint64_t mask);
uint64_t splitmix64(uint64_t &s);
double uniform01(uint64_t &s);
uint32_t density_threshold(double density);
uint32_t render_density_threshold(double density, bool *all_alive);
DevInit make_init(const uint32_t dims[3], const uint32_t spawn[3], uint32_t spawn_mode, double density,
                  uint32_t seed);

}  // namespace cave
