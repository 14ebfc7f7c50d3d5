// Internal declarations shared by the host and HIP halves of libnebula.so.
#pragma once

#include <cstdint>
#include <string>
#include <vector>

#include "../../../include/nebula.h"

namespace nbl {

struct V3 { float x, y, z; };

// Record `msg` as the calling thread's last error and return `code`.
int fail(int code, const std::string& msg);
void clear_error();

int build_octree(const nbl_triangle44* tris, int32_t ntri, std::vector<int32_t>& od,
	std::vector<nbl_triangle44>& sorted, float center[3], float halfsize[3]);
int prepare_material(const nbl_raw_material* r, nbl_material_host** out);
float material_max_energy(const nbl_material_host* m);
int load_tri_file(const char* path, std::vector<nbl_triangle44>& out);
int load_pri_file(const char* path, const float mn[3], const float mx[3], float max_E,
	std::vector<nbl_record36>& out, int64_t warnings[4]);
void prescan_shuffle(nbl_record36* rec, int64_t n, int64_t prescan_size);
void sort_primaries(nbl_record36* rec, int64_t n);

} // namespace nbl
