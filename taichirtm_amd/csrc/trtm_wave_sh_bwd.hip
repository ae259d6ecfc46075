// Instantiations of the one-wavefront-per-tile kernel: SH system, backward steps.
#include "trtm_wave_kernel.hpp"

namespace trtm {

TRTM_WAVE_UNIT(launch_wave_sh_bwd) {
    if (tx == 16 && k == 4) {
        launch_wave_cfg<false, false, true, 16, 4>(m, shots, n_shots, it0, tiles_x, tiles_z, s);
        return true;
    }
    if (tx == 8 && k == 4) {
        launch_wave_cfg<false, false, true, 8, 4>(m, shots, n_shots, it0, tiles_x, tiles_z, s);
        return true;
    }
    if (tx == 24 && k == 4) {
        launch_wave_cfg<false, false, true, 24, 4>(m, shots, n_shots, it0, tiles_x, tiles_z, s);
        return true;
    }
    if (tx == 16 && k == 2) {
        launch_wave_cfg<false, false, true, 16, 2>(m, shots, n_shots, it0, tiles_x, tiles_z, s);
        return true;
    }
    return false;
}

}  // namespace trtm
