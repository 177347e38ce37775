// Writes the image files of 2d_strip_reconstruction through this build's
// writers (j-pet-mlem-old_amd/host/io.hpp) and reads --rho start images, so
// tests/test_io.py can compare them byte for byte with the files the
// reference's own writers produce (tests/golden/io/).
//
//   io_writers write <image.bin> <n_z> <n_y> <out_dir>
//   io_writers read  <start image> <n_z> <n_y> <is_int> <out.bin>
#include <cstdlib>
#include <string>
#include <vector>

#include "io.hpp"

int main(int argc, char** argv) {
  if (argc < 6) return 2;
  const std::string cmd = argv[1];
  const int n_z = std::atoi(argv[3]), n_y = std::atoi(argv[4]);
  const size_t n = (size_t)n_z * n_y;
  std::vector<float> img(n, 1.0f);
  if (cmd == "write") {
    std::ifstream in(argv[2], std::ios::binary);
    in.read(reinterpret_cast<char*>(img.data()), sizeof(float) * n);
    const std::string d = argv[5];
    strip_io::write_matrix_text(d + "/matrix.txt", img.data(), n_z, n_y);
    strip_io::write_tuples_text(d + "/tuples.txt", img.data(), n_z, n_y);
    strip_io::write_raw(d + "/image.bin", img.data(), n);
    strip_io::write_nrrd(d + "/raw.nrrd", "image.bin", n_z, n_y, false);
    strip_io::write_nrrd(d + "/ascii.nrrd", "matrix.txt", n_z, n_y, true);
    return 0;
  }
  if (cmd == "read" && argc >= 7) {
    strip_io::read_image(argv[2], std::atoi(argv[5]) != 0, img.data(), n);
    strip_io::write_raw(argv[6], img.data(), n);
    return 0;
  }
  return 2;
}
