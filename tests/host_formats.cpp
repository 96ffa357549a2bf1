// Writes the driver's debug-dump and cell-potential files from fixed inputs
// into the directory argv[1], for tests/test_host_formats.py to compare with
// the text the reference's writers produce (nbody.cu:68-127,412-697,704-724).
#include <cstdio>
#include <string>
#include <vector>

#include "nbody_io.hpp"

using namespace khrr_host;

int main(int argc, char** argv) {
  if (argc != 2) return 2;
  const std::string dir = argv[1];
  const int nx = 5;
  const double sim_l = 0.5 * nx * 0.2, dx = 0.2;
  std::vector<double> cells(nx * nx * nx);
  for (size_t c = 0; c < cells.size(); ++c) cells[c] = 1.0 / (double)(c + 3) - 0.01 * (double)c;
  std::vector<int> count(nx * nx * nx);
  for (size_t c = 0; c < count.size(); ++c) count[c] = (int)((c * 7) % 5) + (c == 62 ? 1234567 : 0);
  std::vector<double> body = {0.1, -2.5, 3.14159265358979, 1e-9, 12345.678901234567, 0.0};
  bool ok = write_x_slice(dir + "/slice7.txt", "x, phi", nx, sim_l, dx, cells, 7) &&
            write_x_slice(dir + "/slice6.txt", "x, mass", nx, sim_l, dx, cells, 6) &&
            write_index_list(dir + "/list.txt", body) &&
            write_count_slice(dir + "/counts.txt", nx, count) &&
            write_count_distribution(dir, count);
  return ok ? 0 : 1;
}
