// 3DGS PLY files (crates/nano-splat/src/ply.rs:17-256): binary
// little-endian, one `element vertex` of float properties in the Inria
// schema x y z, nx ny nz, f_dc_0..2, f_rest_0..44 (planar R[15] G[15] B[15]),
// opacity (logit), scale_0..2 (log), rot_0..3 (w x y z).  The reader maps
// properties by name, so their order may differ; unknown float properties
// are skipped, missing normals are allowed, missing f_rest_i read as 0, and
// anything else missing is an error.  Host-only code: it runs without a GPU.
#include <stdio.h>
#include <string.h>

#include <string>
#include <vector>

#include "nts_internal.h"

namespace {

struct PlyHeader {
  bool binary_le = false;
  uint64_t vertex_count = 0;
  std::vector<std::string> props;  // all float
  long data_offset = 0;
  int find(const char* name) const {
    for (size_t i = 0; i < props.size(); ++i)
      if (props[i] == name) return (int)i;
    return -1;
  }
};

std::string trim(const std::string& s) {
  size_t a = 0, b = s.size();
  while (a < b && (s[a] == ' ' || s[a] == '\t' || s[a] == '\r' || s[a] == '\n')) ++a;
  while (b > a && (s[b - 1] == ' ' || s[b - 1] == '\t' || s[b - 1] == '\r' || s[b - 1] == '\n')) --b;
  return s.substr(a, b - a);
}

bool starts(const std::string& s, const char* p) { return s.compare(0, strlen(p), p) == 0; }

// ply.rs:151-206
int read_header(FILE* f, PlyHeader& h) {
  char buf[4096];
  if (!fgets(buf, sizeof buf, f) || trim(buf) != "ply") return nts_fail(NTS_E_INVALID, "missing 'ply' magic line");
  for (;;) {
    if (!fgets(buf, sizeof buf, f)) return nts_fail(NTS_E_INVALID, "unterminated PLY header");
    const std::string t = trim(buf);
    if (t.empty() || starts(t, "comment ")) continue;
    if (t == "end_header") break;
    if (starts(t, "format ")) {
      h.binary_le = starts(t.substr(7), "binary_little_endian");
      continue;
    }
    if (starts(t, "element vertex ")) {
      char* end = nullptr;
      const std::string num = trim(t.substr(15));
      const unsigned long long v = strtoull(num.c_str(), &end, 10);
      if (num.empty() || (end && *end)) return nts_fail(NTS_E_INVALID, "element vertex: bad count '%s'", num.c_str());
      h.vertex_count = v;
      continue;
    }
    if (starts(t, "property float ")) {
      h.props.push_back(trim(t.substr(15)));
      continue;
    }
    if (starts(t, "element ")) continue;  // no other elements in 3DGS files
    if (starts(t, "property ")) return nts_fail(NTS_E_INVALID, "only `property float ...` supported, got: %s", t.c_str());
  }
  h.data_offset = ftell(f);
  if (!h.binary_le) return nts_fail(NTS_E_INVALID, "only binary_little_endian PLY is supported");
  return NTS_OK;
}

int open_header(const char* path, FILE** fp, PlyHeader& h) {
  *fp = fopen(path, "rb");
  if (!*fp) return nts_fail(NTS_E_INVALID, "cannot open %s", path);
  const int rc = read_header(*fp, h);
  if (rc) {
    fclose(*fp);
    *fp = nullptr;
  }
  return rc;
}

}  // namespace

extern "C" int nts_ply_count(const char* path, uint32_t* n) {
  NTS_CHECK_ARG(path && n);
  FILE* f = nullptr;
  PlyHeader h;
  int rc = open_header(path, &f, h);
  if (rc) return rc;
  fclose(f);
  if (h.vertex_count > 0xffffffffull) return nts_fail(NTS_E_CAPACITY, "%llu vertices", (unsigned long long)h.vertex_count);
  *n = (uint32_t)h.vertex_count;
  return NTS_OK;
}

// ply.rs:116-134, 208-256
extern "C" int nts_ply_read(const char* path, nts_splats_host* out) {
  NTS_CHECK_ARG(path && out && out->pos && out->rot && out->log_scale && out->logit && out->sh_dc && out->sh_rest);
  FILE* f = nullptr;
  PlyHeader h;
  int rc = open_header(path, &f, h);
  if (rc) return rc;
  struct Closer {
    FILE* f;
    ~Closer() { fclose(f); }
  } closer{f};
  if (h.vertex_count > out->n)
    return nts_fail(NTS_E_CAPACITY, "file holds %llu splats, the arrays %u", (unsigned long long)h.vertex_count,
                    out->n);
  const char* required[] = {"x", "y", "z", "f_dc_0", "f_dc_1", "f_dc_2", "opacity", "scale_0",
                            "scale_1", "scale_2", "rot_0", "rot_1", "rot_2", "rot_3"};
  int idx[14];
  for (int k = 0; k < 14; ++k) {
    idx[k] = h.find(required[k]);
    if (idx[k] < 0) return nts_fail(NTS_E_INVALID, "PLY missing required property: %s", required[k]);
  }
  int rest[45];
  for (int k = 0; k < 45; ++k) {
    char name[16];
    snprintf(name, sizeof name, "f_rest_%d", k);
    rest[k] = h.find(name);
  }
  const size_t rec = h.props.size();
  std::vector<float> buf(rec * 4096);
  uint64_t done = 0;
  while (done < h.vertex_count) {
    const uint64_t take = h.vertex_count - done < 4096 ? h.vertex_count - done : 4096;
    if (fread(buf.data(), 4 * rec, take, f) != take) return nts_fail(NTS_E_INVALID, "PLY data ends early");
    for (uint64_t j = 0; j < take; ++j) {
      const float* r = buf.data() + j * rec;  // little-endian host
      const uint64_t i = done + j;
      for (int k = 0; k < 3; ++k) out->pos[3 * i + k] = r[idx[k]];
      for (int k = 0; k < 3; ++k) out->sh_dc[3 * i + k] = r[idx[3 + k]];
      out->logit[i] = r[idx[6]];
      for (int k = 0; k < 3; ++k) out->log_scale[3 * i + k] = r[idx[7 + k]];
      for (int k = 0; k < 4; ++k) out->rot[4 * i + k] = r[idx[10 + k]];
      for (int k = 0; k < 45; ++k) out->sh_rest[45 * i + k] = rest[k] >= 0 ? r[rest[k]] : 0.0f;
    }
    done += take;
  }
  out->n = (uint32_t)h.vertex_count;
  return NTS_OK;
}

// ply.rs:17-100: the writer's header and record order; normals are (0, 1, 0)
// (the reader's default for files without normals, ply.rs:222-226).
extern "C" int nts_ply_write(const char* path, const nts_splats_host* s) {
  NTS_CHECK_ARG(path && s && (s->n == 0 || (s->pos && s->rot && s->log_scale && s->logit && s->sh_dc && s->sh_rest)));
  FILE* f = fopen(path, "wb");
  if (!f) return nts_fail(NTS_E_INVALID, "cannot create %s", path);
  std::string hdr = "ply\nformat binary_little_endian 1.0\nelement vertex " + std::to_string(s->n) + "\n";
  const char* head[] = {"x", "y", "z", "nx", "ny", "nz", "f_dc_0", "f_dc_1", "f_dc_2"};
  for (const char* p : head) hdr += std::string("property float ") + p + "\n";
  for (int k = 0; k < 45; ++k) hdr += "property float f_rest_" + std::to_string(k) + "\n";
  const char* tail[] = {"opacity", "scale_0", "scale_1", "scale_2", "rot_0", "rot_1", "rot_2", "rot_3"};
  for (const char* p : tail) hdr += std::string("property float ") + p + "\n";
  hdr += "end_header\n";
  bool ok = fwrite(hdr.data(), 1, hdr.size(), f) == hdr.size();
  std::vector<float> rec(62);
  for (uint32_t i = 0; ok && i < s->n; ++i) {
    float* r = rec.data();
    for (int k = 0; k < 3; ++k) r[k] = s->pos[3 * (size_t)i + k];
    r[3] = 0.0f;
    r[4] = 1.0f;
    r[5] = 0.0f;
    for (int k = 0; k < 3; ++k) r[6 + k] = s->sh_dc[3 * (size_t)i + k];
    for (int k = 0; k < 45; ++k) r[9 + k] = s->sh_rest[45 * (size_t)i + k];
    r[54] = s->logit[i];
    for (int k = 0; k < 3; ++k) r[55 + k] = s->log_scale[3 * (size_t)i + k];
    for (int k = 0; k < 4; ++k) r[58 + k] = s->rot[4 * (size_t)i + k];
    ok = fwrite(r, 4, 62, f) == 62;
  }
  if (fclose(f) != 0) ok = false;
  return ok ? NTS_OK : nts_fail(NTS_E_INVALID, "write to %s failed", path);
}
