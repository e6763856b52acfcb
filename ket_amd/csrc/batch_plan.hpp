// Host side of a batched sweep (ket_batch_run) that needs no GPU: the
// per-instance planning, run on up to 16 host threads, and the upload pool's
// layout and parallel write.
#pragma once

#include <algorithm>
#include <atomic>
#include <cstddef>
#include <cstring>
#include <deque>
#include <string>
#include <thread>
#include <vector>

#include "kernels.hpp"
#include "ket_amd.h"
#include "parallel.hpp"
#include "plan.hpp"

namespace ketbatch {

inline std::size_t align(std::size_t x, std::size_t a) { return (x + a - 1) / a * a; }

using ketpar::parallel_for;

// The upload pool of one sweep, laid out first and written once: small
// tables are kept here until the upload; block records are copied straight
// from the instances' plans into the pinned staging buffer, in parallel.
class PoolBuilder {
 public:
  std::size_t size() const { return size_; }
  std::size_t reserve(std::size_t bytes) {
    const std::size_t off = align(size_, 256);
    size_ = off + bytes;
    return off;
  }
  template <class T>
  std::size_t put(const std::vector<T>& v) {
    const std::size_t bytes = v.size() * sizeof(T);
    const std::size_t off = reserve(bytes);
    if (bytes) {
      const auto* p = reinterpret_cast<const unsigned char*>(v.data());
      owned_.emplace_back(p, p + bytes);
      jobs_.push_back({off, owned_.back().data(), bytes});
    }
    return off;
  }
  // `src` must stay valid until write_to().
  void copy_later(std::size_t dst, const void* src, std::size_t bytes) {
    jobs_.push_back({dst, src, bytes});
  }
  void write_to(unsigned char* dst) const {
    const std::size_t per = 64, chunks = (jobs_.size() + per - 1) / per;
    parallel_for(chunks, [&](std::size_t c) {
      for (std::size_t j = c * per; j < std::min(jobs_.size(), (c + 1) * per); ++j)
        std::memcpy(dst + jobs_[j].dst, jobs_[j].src, jobs_[j].n);
    });
  }

 private:
  struct Job {
    std::size_t dst;
    const void* src;
    std::size_t n;
  };
  std::size_t size_ = 0;
  std::vector<Job> jobs_;
  std::deque<std::vector<unsigned char>> owned_;
};

// A per-gate pass in the LDS-resident kernel's form.
ketk::LdsOp lds_op(const ketplan::Op& o);

// Blocks-only plan of one instance (every packed group a fused block).
int plan_blocks_only(const ket_gate* g, std::size_t ng, int n, int precision, int block_k,
                     ketplan::Plan* plan, std::string* err);

// Runs-only plan of one instance: the run packer's reordered passes, every
// pass in an exact run with fused multiply-adds (KET_FUSE_RUNS's form).
int plan_runs_only(const ket_gate* g, std::size_t ng, int n, int precision,
                   ketplan::Plan* plan, std::string* err);

// Every instance's host work, in parallel: the LDS ops of instances with
// in_lds[i] set, the blocks-only (runs = false) or runs-only plans of the
// others. Instance i's gates are gates[gate_off[i] .. gate_off[i+1]). On
// failure returns the first failing instance's status and message.
int plan_instances(const std::vector<int>& n, const std::vector<char>& in_lds,
                   const ket_gate* gates, const uint64_t* gate_off, int precision, int block_k,
                   std::vector<std::vector<ketk::LdsOp>>* lds_ops,
                   std::vector<ketplan::Plan>* plans, std::string* err, bool runs = false);

}  // namespace ketbatch
