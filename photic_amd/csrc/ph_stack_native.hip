// photic_amd -- the stack kernel's native-bottom instances (1-3 bottoms with the reference's
// own amplitudes and weights, model/samodel.c:1808-1815), compiled beside ph_stack.hip.
#include <hip/hip_runtime.h>

#include "ph_stackkern.h"

extern "C" int ph_stack_occupancy_native(int k, int nbot, int block, int* bpc)
{
    return ph_stack_dates(k, [&](auto kd) {
        constexpr int K = decltype(kd)::value;
        switch (nbot) {
        case 1: return ph_stack_occupancy_t<K, 1, false>(block, bpc);
        case 2: return ph_stack_occupancy_t<K, 2, false>(block, bpc);
        case 3: return ph_stack_occupancy_t<K, 3, false>(block, bpc);
        default: return PH_LAUNCH_UNSUPPORTED;
        }
    });
}

extern "C" int ph_launch_stack_native(const PhStackArgs* a, int k, int nbot, int grid,
                                      int block, void* stream)
{
    return ph_stack_dates(k, [&](auto kd) {
        constexpr int K = decltype(kd)::value;
        switch (nbot) {
        case 1: return ph_launch_stack_t<K, 1, false>(a, grid, block, stream);
        case 2: return ph_launch_stack_t<K, 2, false>(a, grid, block, stream);
        case 3: return ph_launch_stack_t<K, 3, false>(a, grid, block, stream);
        default: return PH_LAUNCH_UNSUPPORTED;
        }
    });
}
