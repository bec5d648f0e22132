// Internal declarations shared by the host sources of libsemblance.
#pragma once

#include "semblance.h"

#include <cstdint>
#include <cstdio>
#include <string>
#include <vector>

namespace sb {

// Thread-local error detail behind sb_last_error().
void set_error(const std::string& msg);
int fail(int code, const std::string& msg);

// SU header layout: 240 bytes, native endian (su_trace.hpp:42-124 of the
// reference; standard SEG-Y/SU byte positions).
constexpr int kSuHeaderBytes = 240;
constexpr int kOffCdp = 20;
constexpr int kOffOffset = 36;
constexpr int kOffScalco = 70;
constexpr int kOffSx = 72;
constexpr int kOffSy = 76;
constexpr int kOffGx = 80;
constexpr int kOffGy = 84;
constexpr int kOffNs = 114;
constexpr int kOffDt = 116;
constexpr int kOffTracl = 0;
constexpr int kOffTracr = 4;
constexpr int kOffCdpt = 24;

struct SuHeader {
  uint8_t raw[kSuHeaderBytes];
};

}  // namespace sb

struct sb_su_gather {
  int method = SB_METHOD_CMP;
  int ns = 0;
  int dt_us = 0;
  std::vector<int32_t> cdp_off;          // [ncdp+1]
  std::vector<sb::SuHeader> first_hdr;   // header of each CDP's first trace
  std::vector<sb_trace_geom> geom;       // [ntr]
  std::vector<float> samples;            // [ntr][ns]
  std::vector<int32_t> nbr_off;          // CRS: [ncdp+1]
  std::vector<int32_t> nbr;              // CRS neighbour CDP indices
};
