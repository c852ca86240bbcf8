// Reads a spin file with mag_amd::read_spins and writes it back with
// mag_amd::write_spins (host only: no GPU call is made).
#include <iostream>
#include <string>

#include "mag.h"

int main(int argc, char** argv) {
  if (argc != 7) {
    std::cerr << "usage: Lx Ly Lz Lt in out" << std::endl;
    return 2;
  }
  const std::array<int, 4> d = {std::stoi(argv[1]), std::stoi(argv[2]), std::stoi(argv[3]),
                                std::stoi(argv[4])};
  try {
    mag_amd::DataPatternLexicographical pattern(d);
    std::vector<mag_amd::spin> s = mag_amd::read_spins<mag_amd::spin>(argv[5], pattern);
    mag_amd::write_spins(argv[6], s);
  } catch (const std::exception& e) {
    std::cerr << e.what() << std::endl;
    return 1;
  }
  return 0;
}
