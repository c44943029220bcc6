// Single-rank gather-scatter (ogs_t subset) for the C0 hex path.
#pragma once

#include "core.hpp"

namespace libp {

/* glibc rand() stream (TYPE_3 additive feedback generator r[i] = r[i-3] +
 * r[i-31], Park-Miller seeding, 310 discarded values, output r >> 1). The
 * reference picks the owner copy of every gathered node with rand()
 * (libs/ogs/ogsSetup.cpp:264-275); carrying the stream explicitly keeps the
 * pick independent of anything else in the process that calls rand(). */
class randStream_t {
  uint32_t ring[34];
  int pos = 0;

 public:
  explicit randStream_t(unsigned seed = 1);
  int next();
};

namespace ogs {

enum Op { Add };
enum Transpose { NoTrans, Trans };

class ogs_t {
 public:
  platform_t platform;
  dlong N = 0;           // local (E-vector) length
  dlong Ngather = 0;     // gathered (owned) rows
  hlong NgatherGlobal = 0;
  dlong NlocalT = 0, NlocalP = 0;
  bool unique = true;

  // CSR of the T (all copies) and N (owner copy) gather operators, host and device
  memory<dlong> rowStartsT, colIdsT, rowStartsN, colIdsN;
  memory<dlong> blockRowStartsT, blockRowStartsN;
  dlong NrowBlocksT = 0, NrowBlocksN = 0;
  deviceMemory<dlong> o_rowStartsT, o_colIdsT, o_rowStartsN, o_colIdsN;

  /* ogs_t::Setup (libs/ogs/ogsSetup.cpp:44-190) for kind=Signed on one rank.
   * ids: zero entries are skipped; with unique=true one copy per id group is
   * made the owner and the signs are written back into ids (:137-147). */
  void Setup(dlong N, memory<hlong> ids, bool unique, randStream_t& rng, platform_t& platform);

  template <typename T>
  void Gather(deviceMemory<T> o_gv, deviceMemory<T> o_v, const int k, const Op op, const Transpose trans);
  void Scatter(deviceMemory<dfloat> o_v, deviceMemory<dfloat> o_gv, const int k, const Transpose trans);

  // ogs_t::SetupGlobalToLocalMapping (ogsSetup.cpp:867-886)
  void SetupGlobalToLocalMapping(memory<dlong>& GlobalToLocal) const;

 private:
  void setupRowBlocks();  // ogsOperator.cpp:495-560
};

}  // namespace ogs
}  // namespace libp
