// Reads a link file with mag_amd::read_links and writes it back in the
// lexicographical double format, or as an ILDG file when a tenth argument
// "ildg" is given (host only: no GPU call is made).
#include <iostream>
#include <string>

#include "mag.h"

int main(int argc, char** argv) {
  if (argc != 10 && argc != 11) {
    std::cerr << "usage: group Lx Ly Lz Lt in format precision out [ildg]" << std::endl;
    return 2;
  }
  const bool ildg_out = argc == 11 && std::string(argv[10]) == "ildg";
  const int group = std::stoi(argv[1]);
  const std::array<int, 4> d = {std::stoi(argv[2]), std::stoi(argv[3]), std::stoi(argv[4]),
                                std::stoi(argv[5])};
  try {
    if (group == 2) {
      mag_amd::Data::LatticeData<mag_amd::DataPatternLexicographical, mag_amd::su2> c(d);
      mag_amd::read_links(c, argv[6], argv[7], 0, argv[8]);
      if (ildg_out) mag_amd::write_links_ildg(c, argv[9]);
      else mag_amd::write_links_lexicographical(c, argv[9]);
    } else {
      mag_amd::Data::LatticeData<mag_amd::DataPatternLexicographical, mag_amd::su3> c(d);
      mag_amd::read_links(c, argv[6], argv[7], 0, argv[8]);
      if (ildg_out) mag_amd::write_links_ildg(c, argv[9]);
      else mag_amd::write_links_lexicographical(c, argv[9]);
    }
  } catch (const std::exception& e) {
    std::cerr << e.what() << std::endl;
    return 1;
  }
  return 0;
}
