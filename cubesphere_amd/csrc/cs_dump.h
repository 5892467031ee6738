// cs_dump.h -- the checkpoint file of a state (cs_dump.cpp): a 4096-byte
// header, then the five state arrays (U, ent, Bxf, Byf, Bzf) in the layouts
// of cs_upload_state, each at a 4096-byte aligned offset with its own
// checksum.  Streaming interface shared by the host-array entry points and
// the handle's (which stage through pinned memory, cs_kernels.hip).
#pragma once
#include <stdint.h>

#include <functional>

#include "../../include/cubesphere_hip.h"

namespace cs {

constexpr int kDumpSections = 5;
constexpr uint32_t kDumpVersion = 1;
constexpr uint64_t kDumpHeaderBytes = 4096;

// element counts of the five sections for a configuration
void dump_counts(const cs_config* cfg, uint64_t counts[kDumpSections]);

// Section data moves through `chunk` (chunk_n doubles, a multiple of 4):
// get fills chunk[0, n) with doubles [off, off + n) of section q, put consumes
// them.  A non-zero return ends the transfer with that status.
struct DumpIo {
  double* chunk = nullptr;
  uint64_t chunk_n = 0;
  std::function<int(int q, uint64_t off, uint64_t n, double* chunk)> get;
  std::function<int(int q, uint64_t off, uint64_t n, const double* chunk)> put;
};

// Writes path (through path.tmp renamed into place, so an existing dump is
// replaced only by a complete one).
int dump_write(const char* path, const cs_config* cfg, const cs_dump_meta* meta, DumpIo& io);
// Reads path into io.put after checking its header against cfg's grid
// (ntp, nr, nghost, radial mesh and radii must match); meta receives the
// header.  Every section's checksum is verified after its last chunk.
int dump_read(const char* path, const cs_config* cfg, cs_dump_meta* meta, DumpIo& io);

}  // namespace cs
