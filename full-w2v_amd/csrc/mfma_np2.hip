// Instantiations of the BF16-MFMA kernel for rows of up to 256 elements.
#include "mfma_impl.h"

namespace w2v {

const void* mfma_instance_np2(int ntb, int mt, bool hw, bool pc) {
#define W2V_NT(NTB)                                                                     \
  if (ntb == NTB) {                                                                     \
    if (mt == 1) return !hw ? (const void*)sgns_mfma_kernel<2, NTB, 1, false>         \
                            : pc ? (const void*)sgns_mfma_kernel<2, NTB, 1, true, true> \
                                 : (const void*)sgns_mfma_kernel<2, NTB, 1, true, false>; \
    return !hw ? (const void*)sgns_mfma_kernel<2, NTB, 2, false>                      \
               : pc ? (const void*)sgns_mfma_kernel<2, NTB, 2, true, true>            \
                    : (const void*)sgns_mfma_kernel<2, NTB, 2, true, false>;          \
  }
  W2V_NT(2) W2V_NT(4) W2V_NT(6) W2V_NT(8) W2V_NT(12) W2V_NT(16)
#undef W2V_NT
  return nullptr;
}

}  // namespace w2v
