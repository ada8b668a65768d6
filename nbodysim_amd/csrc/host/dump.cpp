// dump.cpp -- frame files of src/main.cpp:680-811 with the same bytes:
//   NNNN.dat  raw particle records {pos[3], vel[3], acc[3], mass} in the
//             datatype (80 B double / 40 B float), native endian, no header;
//   NNNN.txt  "x, y, z\n" per particle, ostream default format
//             (precision 6, %g style).
// Files are written under a temporary name and renamed into place, so a
// crash never leaves a truncated frame under the final name.
#include <sys/stat.h>
#include <unistd.h>

#include <cstdio>
#include <cstring>
#include <fstream>
#include <sstream>
#include <string>
#include <vector>

#include "../../../include/nbs.h"
#include "../../../include/nbs_host.h"

namespace {

bool exists(const std::string& p) {
  struct stat st;
  return stat(p.c_str(), &st) == 0;
}

template <class T>
int write_records(const char* path, size_t n, const double* pos,
                  const double* vel, const double* acc, const double* mass) {
  const std::string tmp = std::string(path) + ".tmp";
  FILE* f = fopen(tmp.c_str(), "wb");
  if (!f) return NBS_ERR_IO;
  std::vector<T> buf;
  const size_t chunk = 1 << 16;
  buf.resize(chunk * 10);
  for (size_t b = 0; b < n; b += chunk) {
    const size_t e = b + chunk < n ? b + chunk : n;
    for (size_t i = b; i < e; i++) {
      T* r = &buf[(i - b) * 10];
      for (int k = 0; k < 3; k++) {
        r[k] = (T)pos[3 * i + k];
        r[3 + k] = vel ? (T)vel[3 * i + k] : (T)0;
        r[6 + k] = acc ? (T)acc[3 * i + k] : (T)0;
      }
      r[9] = (T)mass[i];
    }
    if (fwrite(buf.data(), sizeof(T) * 10, e - b, f) != e - b) {
      fclose(f);
      unlink(tmp.c_str());
      return NBS_ERR_IO;
    }
  }
  if (fclose(f) != 0) {
    unlink(tmp.c_str());
    return NBS_ERR_IO;
  }
  return rename(tmp.c_str(), path) == 0 ? NBS_OK : NBS_ERR_IO;
}

template <class T>
int64_t read_records(const char* path, size_t cap, double* pos, double* vel,
                     double* acc, double* mass) {
  FILE* f = fopen(path, "rb");
  if (!f) return -1;
  T r[10];
  size_t i = 0;
  while (i < cap && fread(r, sizeof(T) * 10, 1, f) == 1) {
    for (int k = 0; k < 3; k++) {
      if (pos) pos[3 * i + k] = r[k];
      if (vel) vel[3 * i + k] = r[3 + k];
      if (acc) acc[3 * i + k] = r[6 + k];
    }
    if (mass) mass[i] = r[9];
    i++;
  }
  fclose(f);
  return (int64_t)i;
}

}  // namespace

extern "C" {

int nbs_frame_filename(const char* dir, unsigned frame, int binary, char* out,
                       size_t len) {
  // std::to_string(frame) left-padded with zeros to 4 characters.
  int w = snprintf(out, len, "%s/%04u.%s", dir ? dir : ".", frame,
                   binary ? "dat" : "txt");
  return (w < 0 || (size_t)w >= len) ? NBS_ERR_ARG : NBS_OK;
}

int nbs_write_binary(const char* path, size_t n, const double* pos,
                     const double* vel, const double* acc, const double* mass,
                     int prec32) {
  if (!path || (n && (!pos || !mass))) return NBS_ERR_ARG;
  return prec32 ? write_records<float>(path, n, pos, vel, acc, mass)
                : write_records<double>(path, n, pos, vel, acc, mass);
}

int nbs_write_text(const char* path, size_t n, const double* pos) {
  if (!path || (n && !pos)) return NBS_ERR_ARG;
  const std::string tmp = std::string(path) + ".tmp";
  std::ofstream f(tmp, std::ios::out);
  if (!f.is_open()) return NBS_ERR_IO;
  for (size_t i = 0; i < n; i++)
    f << pos[3 * i] << ", " << pos[3 * i + 1] << ", " << pos[3 * i + 2] << "\n";
  f.flush();
  const bool ok = !f.bad();
  f.close();
  if (!ok) {
    unlink(tmp.c_str());
    return NBS_ERR_IO;
  }
  return rename(tmp.c_str(), path) == 0 ? NBS_OK : NBS_ERR_IO;
}

int64_t nbs_binary_count(const char* path, int prec32) {
  struct stat st;
  if (stat(path, &st) != 0) return -1;
  const int64_t rec = prec32 ? 40 : 80;
  return (int64_t)st.st_size / rec;
}

int64_t nbs_read_binary(const char* path, int prec32, size_t cap, double* pos,
                        double* vel, double* acc, double* mass) {
  return prec32 ? read_records<float>(path, cap, pos, vel, acc, mass)
                : read_records<double>(path, cap, pos, vel, acc, mass);
}

int nbs_resume_frame(const char* dir, unsigned num_frames) {
  char name[4096];
  for (long f = num_frames; f >= 0; f--) {
    if (nbs_frame_filename(dir, (unsigned)f, 1, name, sizeof(name))) return 0;
    if (exists(name)) return (int)(f + 1);
  }
  return 0;
}

}  // extern "C"
