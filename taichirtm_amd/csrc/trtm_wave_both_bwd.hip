// Instantiations of the one-wavefront-per-tile kernel: P-SV + SH system, backward steps.
#include "trtm_wave_kernel.hpp"

namespace trtm {

TRTM_WAVE_UNIT(launch_wave_both_bwd) {
    if (tx == 8 && k == 4) {
        launch_wave_cfg<false, true, true, 8, 4>(m, shots, n_shots, it0, tiles_x, tiles_z, s);
        return true;
    }
    if (tx == 8 && k == 2) {
        launch_wave_cfg<false, true, true, 8, 2>(m, shots, n_shots, it0, tiles_x, tiles_z, s);
        return true;
    }
    return false;
}

}  // namespace trtm
