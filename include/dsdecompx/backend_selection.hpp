// backend_selection.hpp -- backend and execution-mode enums (reference
// include/backend_selection.hpp:8-27).  This build has one transport, the
// single-GPU loopback; ExecutionMode::Cuda keeps its name and means "device
// memory, HIP stream".
#pragma once

#include <string_view>

namespace dsdecompx {

enum class BackendKind { Mpi, Nccl, Nvshmem, Loopback };
enum class BackendSelection { Auto, Mpi, Nccl, Nvshmem };
enum class ExecutionMode { Host, Cuda };

const char* backend_selection_name(BackendSelection selection) noexcept;
BackendSelection parse_backend_selection(std::string_view value);
bool backend_compiled(BackendSelection selection) noexcept;

}  // namespace dsdecompx
