// io_driver.cpp -- hmap::Array's interchange formats through the C++ drop-in
// (include/hmap_gpu.hpp), for tests/test_io.py:
//   io_driver write <dir>   writes a 5 x 3 ramp as raw (a.f32) and .npy (a.npy)
//   io_driver read <file.npy> <out.f32>   from_numpy, then to_file
// No device is used.
#include <cstdio>
#include <string>

#include "hmap_gpu.hpp"

int main(int argc, char **argv)
{
  if (argc < 3) return 2;
  const std::string cmd = argv[1];
  try
  {
    if (cmd == "write")
    {
      hmap::Array a(glm::ivec2(5, 3));
      for (int j = 0; j < 3; j++)
        for (int i = 0; i < 5; i++) a(i, j) = (float)(10 * i + j) + 0.25f;
      a.to_file(std::string(argv[2]) + "/a.f32");
      a.to_numpy(std::string(argv[2]) + "/a.npy");
      hmap::Array b(glm::ivec2(5, 3));
      b.from_file(std::string(argv[2]) + "/a.f32");
      for (size_t k = 0; k < a.size(); k++)
        if (a.vector[k] != b.vector[k]) return 3;
      return 0;
    }
    if (cmd == "read" && argc >= 4)
    {
      hmap::Array a;
      a.from_numpy(argv[2]);
      std::printf("%d %d\n", a.shape.x, a.shape.y);
      a.to_file(argv[3]);
      return 0;
    }
  }
  catch (const std::exception &e)
  {
    std::fprintf(stderr, "error: %s\n", e.what());
    return 1;
  }
  return 2;
}
