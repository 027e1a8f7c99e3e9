// Host-side STL I/O and the surface colour map.
#pragma once
#include <string>
#include <vector>

namespace sr {

int stl_read(const std::string &path, std::vector<float> &tris, std::string &err);
int stl_write(const std::string &path, const float *tris, int n, std::string &err);
void colour_map(const double *p, long n, float minDb, float maxDb, bool phase, float *rgba);

}  // namespace sr
