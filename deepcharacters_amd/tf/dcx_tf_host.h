// Host tables for the TensorFlow op shim's constructors, read through the
// embedding Python interpreter: the reference's ops parse character, graph
// and mesh files in their constructors (DQSkinningGPUOp.cpp:22-90,
// EmbeddedGraphGPUOp.cpp:31-80, ProjectedMeshBoundaryGPUOp.cpp); the shim
// hands the same file paths to deepcharacters_amd.tf.character_io, whose
// readers are the ones the rest of the package uses and tests.
#pragma once

#include <stdint.h>

#include <map>
#include <string>
#include <vector>

namespace dcx_tf {

struct HostArray {
  std::vector<char> bytes;
  std::vector<long long> shape;
  char kind = 0;  // 'f' float32, 'i' int32
  long long size() const { return (long long)bytes.size() / 4; }
  const float* f() const { return reinterpret_cast<const float*>(bytes.data()); }
  const int32_t* i() const { return reinterpret_cast<const int32_t*>(bytes.data()); }
};
using HostTables = std::map<std::string, HostArray>;

// Calls deepcharacters_amd.tf.character_io.<fn>(*args) and copies every array
// of the dict it returns.  Needs an initialised Python interpreter (the TF
// Python front end that loaded the op library); takes the GIL itself.
// Returns an empty string on success, else the error text.
std::string CallCharacterIO(const std::string& fn, const std::vector<std::string>& args, HostTables* out);

}  // namespace dcx_tf
