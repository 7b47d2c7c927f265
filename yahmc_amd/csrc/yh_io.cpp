// Gauge configuration files (host side).
//
// native : library/gaugefield.cpp:44-179, include/gaugefield.h:20-36 - a
//          256-byte header_t (ident, version, T, X, Y, Z, NC, repr,
//          num_cnfg, num_accept, plaquette, reserved), then for t, x, y, z
//          in lexicographic order and mu = 0..3 one suNg written element by
//          element as (re, im) doubles (suNg.cpp:392-412), host byte order;
// HiRep  : library/hirep.cpp:56-212 - five big-endian ints (NC, T, X, Y, Z)
//          and a big-endian double plaquette, then the same link order with
//          big-endian doubles.
// Links cross the ABI in the gfield_fund layout (EO site order, re plane
// then im plane per link).
#include <stdio.h>
#include <string.h>
#include <string>
#include <vector>
#include "yh_internal.hpp"

using namespace yh;

namespace {

const int CFG_IDENT = 0x47464359;
const int CFG_VERSION = 0x20;

#pragma pack(push, 1)
struct Header {
  int ident, version, dim_t, dim_x, dim_y, dim_z, nc, repr, num_cnfg, num_accept;
  double plaquette;
  char reserved[208];
};
#pragma pack(pop)
static_assert(sizeof(Header) == 256, "header_t is 256 bytes");

template <typename T>
T bswap(T v) {
  unsigned char b[sizeof(T)];
  memcpy(b, &v, sizeof(T));
  for (size_t i = 0; i < sizeof(T) / 2; i++) std::swap(b[i], b[sizeof(T) - 1 - i]);
  memcpy(&v, b, sizeof(T));
  return v;
}

bool host_big_endian() {
  const unsigned one = 1;
  unsigned char c;
  memcpy(&c, &one, 1);
  return c == 0;
}

// EO index of lexicographic site (t, x, y, z) (geometry.cpp:756-878)
long eo_index(const int* L, int t, int x, int y, int z) {
  const long V = (long)L[0] * L[1] * L[2] * L[3];
  const long lex = (((long)t * L[1] + x) * L[2] + y) * L[3] + z;
  return ((t + x + y + z) & 1) * (V / 2) + lex / 2;
}

struct File {
  FILE* f = nullptr;
  ~File() {
    if (f) fclose(f);
  }
};

}  // namespace

extern "C" {

int yh_config_write(const char* path, int format, const yh_config_info* info, const double* links) {
  if (!path || !info || !links || (format != YH_CFG_NATIVE && format != YH_CFG_HIREP)) {
    set_error("bad argument");
    return YH_ERR_ARG;
  }
  const int* L = info->dims;
  const int nc = info->nc;
  if (nc < 2 || nc > 4 || L[0] < 1 || L[1] < 1 || L[2] < 1 || L[3] < 1 || L[3] % 2) {
    set_error("bad lattice in the configuration info");
    return YH_ERR_ARG;
  }
  File F;
  F.f = fopen(path, "wb");
  if (!F.f) { set_error(std::string("unable to open ") + path); return YH_ERR_ARG; }
  const bool swap = format == YH_CFG_HIREP && !host_big_endian();
  if (format == YH_CFG_NATIVE) {
    Header h;
    memset(&h, 0, sizeof(h));
    h.ident = CFG_IDENT;
    h.version = CFG_VERSION;
    h.dim_t = L[0];
    h.dim_x = L[1];
    h.dim_y = L[2];
    h.dim_z = L[3];
    h.nc = nc;
    h.repr = info->repr;
    h.num_cnfg = info->num_cnfg;
    h.num_accept = info->num_accept;
    h.plaquette = info->plaquette;
    if (fwrite(&h, sizeof(h), 1, F.f) != 1) { set_error("write failed"); return YH_ERR_ARG; }
  } else {
    int head[5] = {nc, L[0], L[1], L[2], L[3]};
    double plaq = info->plaquette;
    if (swap) {
      for (int& v : head) v = bswap(v);
      plaq = bswap(plaq);
    }
    if (fwrite(head, sizeof(int), 5, F.f) != 5 || fwrite(&plaq, sizeof(double), 1, F.f) != 1) {
      set_error("write failed");
      return YH_ERR_ARG;
    }
  }
  const int n2 = nc * nc;
  std::vector<double> buf((size_t)L[3] * 4 * 2 * n2);
  for (int t = 0; t < L[0]; t++)
    for (int x = 0; x < L[1]; x++)
      for (int y = 0; y < L[2]; y++) {
        size_t k = 0;
        for (int z = 0; z < L[3]; z++) {
          const long id = eo_index(L, t, x, y, z);
          for (int mu = 0; mu < 4; mu++) {
            const double* u = links + (size_t)(4 * id + mu) * 2 * n2;
            for (int e = 0; e < n2; e++) {
              buf[k++] = swap ? bswap(u[e]) : u[e];
              buf[k++] = swap ? bswap(u[n2 + e]) : u[n2 + e];
            }
          }
        }
        if (fwrite(buf.data(), sizeof(double), k, F.f) != k) { set_error("write failed"); return YH_ERR_ARG; }
      }
  return YH_OK;
}

int yh_config_read(const char* path, int format, yh_config_info* info, double* links) {
  if (!path || !info || (format != YH_CFG_NATIVE && format != YH_CFG_HIREP)) {
    set_error("bad argument");
    return YH_ERR_ARG;
  }
  File F;
  F.f = fopen(path, "rb");
  if (!F.f) { set_error(std::string("unable to open ") + path); return YH_ERR_ARG; }
  const bool swap = format == YH_CFG_HIREP && !host_big_endian();
  int L[4], nc;
  if (format == YH_CFG_NATIVE) {
    Header h;
    if (fread(&h, sizeof(h), 1, F.f) != 1) { set_error("short header"); return YH_ERR_ARG; }
    if (h.ident != CFG_IDENT || h.version != CFG_VERSION) {
      set_error("not a configuration file of this format (ident / version)");
      return YH_ERR_ARG;
    }
    L[0] = h.dim_t; L[1] = h.dim_x; L[2] = h.dim_y; L[3] = h.dim_z;
    nc = h.nc;
    info->repr = h.repr;
    info->num_cnfg = h.num_cnfg;
    info->num_accept = h.num_accept;
    info->plaquette = h.plaquette;
  } else {
    int head[5];
    double plaq;
    if (fread(head, sizeof(int), 5, F.f) != 5 || fread(&plaq, sizeof(double), 1, F.f) != 1) {
      set_error("short header");
      return YH_ERR_ARG;
    }
    if (swap) {
      for (int& v : head) v = bswap(v);
      plaq = bswap(plaq);
    }
    nc = head[0];
    for (int i = 0; i < 4; i++) L[i] = head[1 + i];
    info->repr = -1;
    info->num_accept = 0;
    info->plaquette = plaq;
    // the trajectory number follows the last 'n' of the file name (hirep.cpp:86-92)
    info->num_cnfg = 0;
    const char* p = strrchr(path, 'n');
    if (p) sscanf(p, "n%d", &info->num_cnfg);
  }
  if (nc < 2 || nc > 4 || L[0] < 1 || L[1] < 1 || L[2] < 1 || L[3] < 1 || L[3] % 2) {
    set_error("inconsistent configuration header");
    return YH_ERR_ARG;
  }
  info->nc = nc;
  for (int i = 0; i < 4; i++) info->dims[i] = L[i];
  info->ident = format == YH_CFG_NATIVE ? CFG_IDENT : 0;
  info->version = format == YH_CFG_NATIVE ? CFG_VERSION : 0;
  if (!links) return YH_OK;  // header only
  const int n2 = nc * nc;
  std::vector<double> buf((size_t)L[3] * 4 * 2 * n2);
  for (int t = 0; t < L[0]; t++)
    for (int x = 0; x < L[1]; x++)
      for (int y = 0; y < L[2]; y++) {
        if (fread(buf.data(), sizeof(double), buf.size(), F.f) != buf.size()) {
          set_error("configuration file is truncated");
          return YH_ERR_ARG;
        }
        size_t k = 0;
        for (int z = 0; z < L[3]; z++) {
          const long id = eo_index(L, t, x, y, z);
          for (int mu = 0; mu < 4; mu++) {
            double* u = links + (size_t)(4 * id + mu) * 2 * n2;
            for (int e = 0; e < n2; e++) {
              u[e] = swap ? bswap(buf[k]) : buf[k];
              k++;
              u[n2 + e] = swap ? bswap(buf[k]) : buf[k];
              k++;
            }
          }
        }
      }
  return YH_OK;
}

}  // extern "C"
