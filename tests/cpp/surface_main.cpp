// CPU marching cubes through the source-compatible surface:: header:
// surface::parameterise of a sphere |p - c| = r on an S^3 lattice over
// [0,1]^3; writes the triangles (geometry::MeshTriangle<float> records).
// usage: surface_main S r OUT.bin
#include <cstdlib>
#include <fstream>
#include <iostream>

#include "fluid/surface.hpp"

int main(int argc, char** argv) {
  if (argc < 4) return 2;
  const size_t S = size_t(std::atoi(argv[1]));
  const float r = float(std::atof(argv[2]));
  const auto lattice = surface::createLattice<float>(S, S, S, 0.0f, 1.0f / float(S - 1));
  const auto tris = surface::parameterise<float>(
      0.0f, lattice, [r](glm::tvec3<float>& p) {
        return r - glm::length(p - glm::tvec3<float>(0.5f));
      });
  std::ofstream(argv[3], std::ios::binary)
      .write(reinterpret_cast<const char*>(tris.data()), std::streamsize(tris.size() * sizeof(tris[0])));
  std::cout << "triangles " << tris.size() << std::endl;
  return 0;
}
