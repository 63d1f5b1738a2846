// config.hpp -- Decomp configuration (reference include/config.hpp:11-63).
#pragma once

#include <cstddef>

#include "backend_selection.hpp"
#include "domain.hpp"
#include "grid.hpp"

namespace dsdecompx {

constexpr int kCudaDeviceCurrent = -1;
constexpr int kCudaDeviceByLocalRank = -2;

struct AdvancedBackendConfig {
    BackendSelection selection = BackendSelection::Auto;
    bool read_environment = true;
    bool allow_host_staging = true;
    // GPU of this process: an ordinal, or one of the constants above.
    int cuda_device = kCudaDeviceCurrent;
};

struct DynamicLoadBalanceConfig {
    std::size_t interval = 1;
    double imbalance_threshold = 1.10;
    double relaxation = 0.5;
    double minimum_cell_fraction = 0.10;
};

struct Config {
    GlobalDomain domain;
    ProcessGrid process_grid;
    HaloConfig halo;
    GridPolicy grid_policy = GridPolicy::Manual;
    TopologyInfo topology_info;
    DynamicLoadBalanceConfig dynamic;
    ExecutionMode execution = ExecutionMode::Cuda;
    AdvancedBackendConfig advanced;
};

struct MigrationResult {
    bool ownership_changed = false;
    std::size_t previous_owned_count = 0;
    std::size_t retained_count = 0;
    std::size_t sent_count = 0;
    std::size_t received_count = 0;
    std::size_t new_owned_count = 0;
};

}  // namespace dsdecompx
