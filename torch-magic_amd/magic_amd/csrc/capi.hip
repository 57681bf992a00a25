// ABI self-description: sizes of the argument structs as compiled, so a
// binding can check its own struct layouts against the library's.
#include "common.h"

extern "C" int mh_struct_sizes(int* out, int n) {
  const int sizes[] = {
      (int)sizeof(MhSynArgs),  (int)sizeof(MhAnaArgs),  (int)sizeof(MhFftArgs),
      (int)sizeof(MhNlArgs),   (int)sizeof(MhRowJob),   (int)sizeof(MhTdArgs),
      (int)sizeof(MhRhsArgs),  (int)sizeof(MhBuoyArgs), (int)sizeof(MhImplArgs),
      (int)sizeof(MhEnergyArgs), (int)sizeof(MhBuildArgs), (int)sizeof(MhGemmArgs),
      (int)sizeof(MhGridArgs), (int)sizeof(MhStepState), (int)sizeof(MhStepCtlArgs),
      (int)sizeof(MhIcRhsArgs), (int)sizeof(MhIcImplArgs), (int)sizeof(MhIcRotArgs),
      (int)sizeof(MhNlAnelArgs), (int)sizeof(MhDphiArgs), (int)sizeof(MhEntropyAnelArgs),
      (int)sizeof(MhImplAnelArgs), (int)sizeof(MhAmArgs), (int)sizeof(MhP0IntArgs),
      (int)sizeof(MhDcArgs),
  };
  const int ns = (int)(sizeof(sizes) / sizeof(sizes[0]));
  for (int i = 0; i < n && i < ns; ++i) out[i] = sizes[i];
  return ns;
}
