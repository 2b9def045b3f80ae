// C++ call surface of the tabulated pattern, host only: Preprocessor::ReadPatternFile (the reference's
// TXPattern.dat reader, Preprocessor.cpp:3-35) and RTSolver::SetTxPattern.  Usage:
//   pattern_surface_check <pattern file>
// Prints "<read ok> <entries> <SetTxPattern status> <antennaType> <status without one entry>
// <status of an empty map> <read of a missing file>", then the physical-mode switches of a default
// RTSolverOptions and of one after SetPhysical().
#include <cstdio>
#include <iterator>

#include "sbrsolver/Preprocessor.h"
#include "sbrsolver/RTSolver.h"

int main(int argc, char** argv) {
  if (argc < 2) return 2;
  GainMap m;
  const bool ok = Preprocessor::ReadPatternFile(argv[1], m);
  RTSolver s;
  const int rc = s.SetTxPattern(m);
  const int type = s.options.antennaType;
  GainMap holed = m;
  if (!holed.empty()) holed.erase(std::next(holed.begin(), (long)holed.size() / 2));
  RTSolver s2;
  const int rcHoled = s2.SetTxPattern(holed);
  GainMap none;
  const int rcEmpty = s2.SetTxPattern(none);
  GainMap missing;
  const bool okMissing = Preprocessor::ReadPatternFile(std::string(argv[1]) + ".does-not-exist", missing);
  std::printf("%d %zu %d %d %d %d %d\n", ok ? 1 : 0, m.size(), rc, type, rcHoled, rcEmpty, okMissing ? 1 : 0);
  // Second line: the physical-mode switches before and after RTSolverOptions::SetPhysical().
  RTSolverOptions o;
  std::printf("%d %d %d %d %d %d", o.q1MeshIndexing, o.receptionModel, o.typedDedup, o.checkOcclusion, o.longMiss,
              o.physicalField);
  o.SetPhysical();
  std::printf(" %d %d %d %d %d %d %g\n", o.q1MeshIndexing, o.receptionModel, o.typedDedup, o.checkOcclusion,
              o.longMiss, o.physicalField, o.wallWidth);
  return 0;
}
