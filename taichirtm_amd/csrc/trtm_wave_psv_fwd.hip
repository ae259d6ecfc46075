// Instantiations of the one-wavefront-per-tile kernel: P-SV system, forward steps.
#include "trtm_wave_kernel.hpp"

namespace trtm {

TRTM_WAVE_UNIT(launch_wave_psv_fwd) {
    if (tx == 16 && k == 4) {
        launch_wave_cfg<true, true, false, 16, 4>(m, shots, n_shots, it0, tiles_x, tiles_z, s);
        return true;
    }
    if (tx == 8 && k == 4) {
        launch_wave_cfg<true, true, false, 8, 4>(m, shots, n_shots, it0, tiles_x, tiles_z, s);
        return true;
    }
    if (tx == 16 && k == 2) {
        launch_wave_cfg<true, true, false, 16, 2>(m, shots, n_shots, it0, tiles_x, tiles_z, s);
        return true;
    }
    if (tx == 8 && k == 2) {
        launch_wave_cfg<true, true, false, 8, 2>(m, shots, n_shots, it0, tiles_x, tiles_z, s);
        return true;
    }
    return false;
}

}  // namespace trtm
