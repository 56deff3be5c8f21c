// Drives hip::SphSolver (include/fluid/hipsph.hpp) the way libfluid's sample
// application drives ocl::SphSolver (samples/main.cpp:398-442): read
// particles, call advance() `steps` times, write particles and the last
// step's triangles and query results.
//
// usage: dropin_main IN.bin OUT_PARTICLES.bin OUT_TRIANGLES.bin OUT_QUERIES.txt
//        h dt scale resolution isolevel iterations gx gy gz
//        minx miny minz maxx maxy maxz steps [qx qy qz]...
// IN.bin / OUT_PARTICLES.bin: raw fluid::Particle<size_t,float> records.
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <iostream>
#include <vector>

#include "fluid/hipsph.hpp"

int main(int argc, char** argv) {
  if (argc < 21) {
    std::cerr << "usage: see source\n";
    return 2;
  }
  std::ifstream in(argv[1], std::ios::binary | std::ios::ate);
  const size_t bytes = size_t(in.tellg());
  in.seekg(0);
  std::vector<fluid::Particle<size_t, float>> xs(bytes / sizeof(fluid::Particle<size_t, float>));
  in.read(reinterpret_cast<char*>(xs.data()), std::streamsize(bytes));
  auto f = [&](int i) { return float(std::atof(argv[i])); };
  const float h = f(5);
  std::vector<fluid::Query<float>> queries;
  for (int i = 21, id = 0; i + 2 < argc; i += 3, ++id)
    queries.push_back(fluid::Query<float>{size_t(id), glm::tvec3<float>(f(i), f(i + 1), f(i + 2))});
  const fluid::Config<float> config(f(6), f(7), f(8), f(9), size_t(std::atoi(argv[10])),
                                    glm::tvec3<float>(f(11), f(12), f(13)), {}, {}, {}, queries,
                                    glm::tvec3<float>(f(14), f(15), f(16)),
                                    glm::tvec3<float>(f(17), f(18), f(19)));
  const int steps = std::atoi(argv[20]);
  try {
    hip::SphSolver solver(h);
    fluid::Result<float> result;
    for (int s = 0; s < steps; ++s) result = solver.advance(config, xs, {});
    std::ofstream(argv[2], std::ios::binary)
        .write(reinterpret_cast<const char*>(xs.data()), std::streamsize(xs.size() * sizeof(xs[0])));
    std::ofstream(argv[3], std::ios::binary)
        .write(reinterpret_cast<const char*>(result.triangles.data()),
               std::streamsize(result.triangles.size() * sizeof(result.triangles[0])));
    std::ofstream q(argv[4]);
    for (const auto& r : result.queries)
      q << r.id << " " << r.neighbours << " " << r.averageColour << "\n";
    std::cout << "particles " << xs.size() << " triangles " << result.triangles.size()
              << " queries " << result.queries.size() << std::endl;
  } catch (const std::exception& e) {
    std::cerr << "error: " << e.what() << std::endl;
    return 1;
  }
  return 0;
}
