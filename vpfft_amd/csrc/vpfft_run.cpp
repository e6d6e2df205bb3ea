// vpfft_run -- run the strain steps an input deck describes on the GPU.
//
//   vpfft_run <deck> [--mixed] [--stress-strain FILE] [--write-material FILE]
//
// The deck is the reference's ConfigFile format (Src/Config.cpp:37-261);
// Stress_Strain.txt follows MaterialGrid::RunVPFFT (Src/MaterialGrid.cpp:1202-1228).
#include <cstdio>
#include <cstring>
#include <vector>

#include "../../include/vpfft.h"
#include "../../include/vpfft_io.h"

int main(int argc, char** argv) {
  const char* deck = nullptr;
  const char* stress_strain = "Stress_Strain.txt";
  const char* material_out = nullptr;
  int precision = VPFFT_FP64;
  for (int i = 1; i < argc; i++) {
    if (!std::strcmp(argv[i], "--mixed")) {
      precision = VPFFT_MIXED;
    } else if (!std::strcmp(argv[i], "--stress-strain") && i + 1 < argc) {
      stress_strain = argv[++i];
    } else if (!std::strcmp(argv[i], "--write-material") && i + 1 < argc) {
      material_out = argv[++i];
    } else if (argv[i][0] != '-' && !deck) {
      deck = argv[i];
    } else {
      deck = nullptr;
      break;
    }
  }
  if (!deck) {
    std::fprintf(stderr, "usage: %s <deck> [--mixed] [--stress-strain FILE] [--write-material FILE]\n", argv[0]);
    return 2;
  }
  vpfft_deck d;
  if (vpfft_deck_read(deck, &d)) {
    std::fprintf(stderr, "vpfft_run: %s\n", vpfft_last_error());
    return 1;
  }
  std::printf("GridDimension %d %d %d, %d steps of dt %g, at most %d iterations, tolerance %g\n", d.num_x, d.num_y,
              d.num_z, d.num_time_iter, d.time_step, d.max_nr_iter, d.nr_tolerance);
  std::vector<vpfft_step_info> info(d.num_time_iter > 0 ? d.num_time_iter : 1);
  if (vpfft_run_deck(deck, precision, stress_strain, material_out, info.data(), (int)info.size())) {
    std::fprintf(stderr, "vpfft_run: %s\n", vpfft_last_error());
    return 1;
  }
  for (int s = 0; s < d.num_time_iter; s++)
    std::printf("step %d: %d iterations, err_stress %.3e, err_strain_rate %.3e, |<sigma>| components %g %g %g %g %g\n",
                s, info[s].iterations, info[s].err_stress, info[s].err_strain_rate, info[s].avg_stress[0],
                info[s].avg_stress[1], info[s].avg_stress[2], info[s].avg_stress[3], info[s].avg_stress[4]);
  return 0;
}
