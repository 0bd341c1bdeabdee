// io.cpp -- the reference's file formats for the path's inputs and outputs
// (lib/src/qcd_io.c), on one rank, written from the format definitions:
//
// * parameter files: the whole file is read; "<tag>" is found by byte
//   comparison and the value is the text after it up to the next '<'
//   (qcd_io.c:83-128);
// * LIME files: a sequence of records, each a 144-byte big-endian header
//   (magic 1164413355 as 32 bits, version 16 bits, a byte with the
//   message-begin 0x80 / message-end 0x40 flags, one reserved byte, the data
//   length as 64 bits, then a 128-byte NUL-padded record type string,
//   include/lime_binary_header.h) followed by the data padded to 8 bytes;
// * ILDG gauge fields: record "ildg-format" holds <precision> (64 or 32),
//   "ildg-binary-data" big-endian numbers in the order t, z, y, x, mu', row,
//   column, re/im, where the file's mu' = 0..3 is x, y, z, t, i.e. our
//   direction (mu' + 1) % 4 (qcd_io.c:204-399, :421-577);
// * LIME spinors: "propagator-type"/"source-type", then
//   "etmc-propagator-format"/"etmc-source-format" with <precision>, then
//   "scidac-binary-data" in the order t, z, y, x, spin, colour, re/im
//   (qcd_io.c:984-1197, :1766-1950);
// * HMC propagators: a text line "beta kappa mu L T", then 12 big-endian
//   FP64 vectors (source spin nu, colour c2 at index 3 nu + c2), each in the
//   order x, y, z, t, spin, colour, re/im (qcd_io.c:1200-1292).
// Host code only: no device call is made here.
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

extern "C" {
#include "qcd.h"
#include "qcd_io.h"
}

namespace qcdhip {
// dropin.cpp: refreshes the resident copy of a bound gauge field after its
// host data changed
int dropin_gauge_written(const void *host);
}

namespace {

constexpr uint32_t kLimeMagic = 1164413355u;
constexpr size_t kLimeHeader = 144;

uint64_t be_u64(const unsigned char *p) {
  uint64_t v = 0;
  for (int i = 0; i < 8; i++) v = (v << 8) | p[i];
  return v;
}
uint32_t be_u32(const unsigned char *p) { return ((uint32_t)p[0] << 24) | ((uint32_t)p[1] << 16) | ((uint32_t)p[2] << 8) | p[3]; }
void put_be_u64(unsigned char *p, uint64_t v) {
  for (int i = 7; i >= 0; i--) {
    p[i] = (unsigned char)(v & 0xff);
    v >>= 8;
  }
}
void put_be_u32(unsigned char *p, uint32_t v) {
  for (int i = 3; i >= 0; i--) {
    p[i] = (unsigned char)(v & 0xff);
    v >>= 8;
  }
}

// big-endian IEEE numbers <-> host doubles
double be_f64(const unsigned char *p) {
  uint64_t b = be_u64(p);
  double d;
  std::memcpy(&d, &b, 8);
  return d;
}
double be_f32(const unsigned char *p) {
  uint32_t b = be_u32(p);
  float f;
  std::memcpy(&f, &b, 4);
  return (double)f;
}
void put_be_f64(unsigned char *p, double d) {
  uint64_t b;
  std::memcpy(&b, &d, 8);
  put_be_u64(p, b);
}

struct LimeRecord {
  std::string type;
  uint64_t bytes;
  long offset;   // file offset of the data
};

// every record of a LIME file; false (with a message) if it is not one
bool lime_records(FILE *f, const char *fname, const char *who, std::vector<LimeRecord> *out) {
  unsigned char h[kLimeHeader];
  out->clear();
  for (;;) {
    const size_t got = std::fread(h, 1, kLimeHeader, f);
    if (got == 0) return true;
    if (got != kLimeHeader || be_u32(h) != kLimeMagic) {
      std::fprintf(stderr, "Error in %s! %s is not a LIME file (bad record header)\n", who, fname);
      return false;
    }
    LimeRecord r;
    r.bytes = be_u64(h + 8);
    char type[129];
    std::memcpy(type, h + 16, 128);
    type[128] = '\0';
    r.type = type;
    r.offset = std::ftell(f);
    out->push_back(r);
    const uint64_t padded = (r.bytes + 7) / 8 * 8;
    if (std::fseek(f, (long)padded, SEEK_CUR) != 0) {
      std::fprintf(stderr, "Error in %s! %s ends inside a record\n", who, fname);
      return false;
    }
  }
}

std::string read_record(FILE *f, const LimeRecord &r) {
  std::string s(r.bytes, '\0');
  std::fseek(f, r.offset, SEEK_SET);
  if (r.bytes && std::fread(&s[0], 1, r.bytes, f) != r.bytes) s.clear();
  return s;
}

// <precision> of an XML record: 64, 32 or 0
int xml_precision(const std::string &xml) {
  const char *p = std::strstr(xml.c_str(), "<precision>");
  if (!p) return 0;
  return std::atoi(p + std::strlen("<precision>"));
}

struct LimeWriter {
  FILE *f;
  int n = 0, total;
  bool ok = true;
  LimeWriter(FILE *file, int records) : f(file), total(records) {}
  void record(const char *type, const void *data, uint64_t bytes) {
    unsigned char h[kLimeHeader];
    std::memset(h, 0, sizeof(h));
    put_be_u32(h, kLimeMagic);
    h[4] = 0;
    h[5] = 1;   // version 1
    h[6] = (unsigned char)((n == 0 ? 0x80 : 0) | (n == total - 1 ? 0x40 : 0));
    put_be_u64(h + 8, bytes);
    std::strncpy((char *)h + 16, type, 127);
    ok = ok && std::fwrite(h, 1, kLimeHeader, f) == kLimeHeader;
    ok = ok && (bytes == 0 || std::fwrite(data, 1, bytes, f) == bytes);
    static const unsigned char zeros[8] = {0};
    const size_t pad = (size_t)((8 - bytes % 8) % 8);
    ok = ok && (pad == 0 || std::fwrite(zeros, 1, pad, f) == pad);
    n++;
  }
};

inline unsigned long lexic(int t, int x, int y, int z, const qcd_uint_2 L[4]) {
  return (unsigned long)t + L[0] * ((unsigned long)x + L[1] * ((unsigned long)y + L[2] * (unsigned long)z));
}

}  // namespace

extern "C" {

// ------------------------------------------------ byte order (qcd_io.c:20-75)
int qcd_isBigEndian(void) {
  const uint32_t one = 1;
  unsigned char b[4];
  std::memcpy(b, &one, 4);
  return b[3] == 1 ? 1 : b[0] == 1 ? 0 : -1;
}

void qcd_swap_8(double *Rd, int N) {
  unsigned char *p = (unsigned char *)Rd;
  for (long i = 0; i < (long)N; i++, p += 8)
    for (int k = 0; k < 4; k++) {
      const unsigned char t = p[k];
      p[k] = p[7 - k];
      p[7 - k] = t;
    }
}

void qcd_swap_4(float *Rd, int N) {
  unsigned char *p = (unsigned char *)Rd;
  for (long i = 0; i < (long)N; i++, p += 4)
    for (int k = 0; k < 2; k++) {
      const unsigned char t = p[k];
      p[k] = p[3 - k];
      p[3 - k] = t;
    }
}

// ------------------------------------------------ parameter files (qcd_io.c:83-128)
// The buffer holds the file's bytes plus a terminating NUL (len counts it).
char *qcd_getParams(char *fname, int *len) {
  FILE *f = std::fopen(fname, "r");
  if (!f) {
    std::fprintf(stderr, "Error, cannot open %s for reading\n", fname);
    return nullptr;
  }
  std::fseek(f, 0, SEEK_END);
  const long n = std::ftell(f);
  std::rewind(f);
  char *p = (char *)std::malloc((size_t)n + 1);
  if (!p) {
    std::fclose(f);
    return nullptr;
  }
  const size_t got = std::fread(p, 1, (size_t)n, f);
  p[got] = '\0';
  std::fclose(f);
  *len = (int)got + 1;
  return p;
}

// Value of token: the text after it up to the next '<', which is overwritten
// with a NUL as the reference does.  A missing token gives "" (the reference
// returns a pointer near the end of the buffer).
char *qcd_getParam(char token[], char *params, int len) {
  static char empty[1] = {'\0'};
  const int tl = (int)std::strlen(token);
  for (int i = 0; i + tl < len; i++)
    if (std::memcmp(token, params + i, tl) == 0) {
      char *v = params + i + tl;
      char *lt = std::strchr(v, '<');
      if (lt) *lt = '\0';
      return v;
    }
  empty[0] = '\0';
  return empty;
}

// ------------------------------------------------ gauge fields
int qcd_getLimeMessage(char *fname, qcd_geometry *geo, char **message_ptr) {
  if (!geo || !geo->initialized) {
    std::fprintf(stderr, "Error in qcd_getLimeMessage! Geometry must be properly initialized.\n");
    return 1;
  }
  FILE *f = std::fopen(fname, "rb");
  if (!f) {
    std::fprintf(stderr, "process 0: Error in qcd_getLimeMessage! Could not open %s for reading\n", fname);
    return 1;
  }
  std::vector<LimeRecord> recs;
  if (!lime_records(f, fname, "qcd_getLimeMessage", &recs)) {
    std::fclose(f);
    return 1;
  }
  std::string msg;
  for (const auto &r : recs) {
    if (r.type == "ildg-binary-data") break;
    if (r.type == "xlf-info") msg = read_record(f, r);
  }
  std::fclose(f);
  char *m = (char *)std::malloc(msg.size() + 1);
  std::memcpy(m, msg.c_str(), msg.size() + 1);
  *message_ptr = m;
  return 0;
}

int qcd_getGaugeLime(char *fname, qcd_gaugeField *u) {
  if (!u || !u->initialized) {
    std::fprintf(stderr, "Error in qcd_getGaugeLime! Gauge field must be properly initialized.\n");
    return 1;
  }
  FILE *f = std::fopen(fname, "rb");
  if (!f) {
    std::fprintf(stderr, "process 0: Error in qcd_getGaugeLime! Could not open %s for reading\n", fname);
    return 1;
  }
  std::vector<LimeRecord> recs;
  if (!lime_records(f, fname, "qcd_getGaugeLime", &recs)) {
    std::fclose(f);
    return 1;
  }
  int prec = 0;
  const LimeRecord *bin = nullptr;
  for (const auto &r : recs) {
    if (r.type == "ildg-format") prec = xml_precision(read_record(f, r));
    if (r.type == "ildg-binary-data") {
      bin = &r;
      break;
    }
  }
  const qcd_uint_2 *L = u->geo->lL;
  const unsigned long V = u->geo->lV;
  const size_t w = prec == 64 ? 8 : 4;
  if ((prec != 64 && prec != 32) || !bin || bin->bytes < V * 72 * w) {
    std::fprintf(stderr, "process 0: Error in qcd_getGaugeLime! Unsupported precision or missing/short binary data in %s\n",
                 fname);
    std::fclose(f);
    return 1;
  }
  std::vector<unsigned char> buf((size_t)V * 72 * w);
  std::fseek(f, bin->offset, SEEK_SET);
  const bool ok = std::fread(buf.data(), 1, buf.size(), f) == buf.size();
  std::fclose(f);
  if (!ok) {
    std::fprintf(stderr, "process 0: Error in qcd_getGaugeLime! Short read of %s\n", fname);
    return 1;
  }
  const unsigned char *p = buf.data();
  for (int t = 0; t < L[0]; t++)
    for (int z = 0; z < L[3]; z++)
      for (int y = 0; y < L[2]; y++)
        for (int x = 0; x < L[1]; x++)
          for (int mu = 0; mu < 4; mu++) {
            double *d = &u->D[lexic(t, x, y, z, L)][(mu + 1) % 4][0][0].re;
            for (int k = 0; k < 18; k++, p += w) d[k] = w == 8 ? be_f64(p) : be_f32(p);
          }
  if (qcdhip::dropin_gauge_written(u->D)) {
    std::fprintf(stderr, "process 0: Error in qcd_getGaugeLime! Could not refresh the bound device copy\n");
    return 1;
  }
  return 0;
}

int qcd_getGaugeField(char *fname, int type, qcd_gaugeField *u) {
  if (type == qcd_GF_LIME) return qcd_getGaugeLime(fname, u);
  std::fprintf(stderr, "Error in qcd_getGaugeField! Unknown type.\n");
  return 1;
}

int qcd_writeGaugeLime(char *fname, qcd_gaugeField *u, char *message) {
  if (!u || !u->initialized) {
    std::fprintf(stderr, "Error in qcd_writeGaugeLime! Gauge field must be properly initialized.\n");
    return 1;
  }
  const qcd_uint_2 *L = u->geo->lL;
  char xml[1024];
  std::snprintf(xml, sizeof(xml),
                "<?xml version=\"1.0\" encoding=\"UTF-8\"?>\n<ildgFormat>\n  <version> 1.0 </version>\n"
                "  <field> su3gauge </field>\n  <precision> 64 </precision>\n  <lx> %d </lx>\n  <ly> %d </ly>\n"
                "  <lz> %d </lz>\n  <lt> %d </lt>\n</ildgFormat>",
                L[1], L[2], L[3], L[0]);
  const unsigned long V = u->geo->lV;
  std::vector<unsigned char> buf((size_t)V * 72 * 8);
  unsigned char *p = buf.data();
  for (int t = 0; t < L[0]; t++)
    for (int z = 0; z < L[3]; z++)
      for (int y = 0; y < L[2]; y++)
        for (int x = 0; x < L[1]; x++)
          for (int mu = 0; mu < 4; mu++) {
            const double *d = &u->D[lexic(t, x, y, z, L)][(mu + 1) % 4][0][0].re;
            for (int k = 0; k < 18; k++, p += 8) put_be_f64(p, d[k]);
          }
  FILE *f = std::fopen(fname, "wb");
  if (!f) {
    std::fprintf(stderr, "Error in qcd_writeGaugeLime! Could not open %s for writing\n", fname);
    return 1;
  }
  const char *msg = message ? message : "";
  LimeWriter w(f, 3);
  w.record("ildg-format", xml, std::strlen(xml));
  w.record("xlf-info", msg, std::strlen(msg));
  w.record("ildg-binary-data", buf.data(), buf.size());
  const bool ok = w.ok && std::fclose(f) == 0;
  if (!ok) std::fprintf(stderr, "Error in qcd_writeGaugeLime! Write error on %s\n", fname);
  return ok ? 0 : 1;
}

// ------------------------------------------------ spinor fields
int qcd_getVectorLime(char *fname, qcd_vector *v) {
  if (!v || !v->initialized) {
    std::fprintf(stderr, "Error in qcd_getVectorLime! Vector must be properly initialized.\n");
    return 1;
  }
  FILE *f = std::fopen(fname, "rb");
  if (!f) {
    std::fprintf(stderr, "process 0: Error in qcd_getVectorLime Could not open %s for reading\n", fname);
    return 1;
  }
  std::vector<LimeRecord> recs;
  if (!lime_records(f, fname, "qcd_getVectorLime", &recs)) {
    std::fclose(f);
    return 1;
  }
  int prec = 0;
  bool after_format = false;
  const LimeRecord *bin = nullptr;
  for (const auto &r : recs) {
    if (r.type == "propagator-type") {
      const std::string t = read_record(f, r);
      if (t.compare(0, 30, "DiracFermion_Source_Sink_Pairs") != 0 && t.compare(0, 17, "DiracFermion_Sink") != 0) {
        std::fprintf(stderr, " process 0: Error in qcd_getVectorLime! Got %s for \"propagator-type\", expecting %s or %s\n",
                     t.c_str(), "DiracFermion_Source_Sink_Pairs", "DiracFermion_Sink");
        std::fclose(f);
        return 1;
      }
    }
    if (r.type == "etmc-propagator-format" || r.type == "etmc-source-format") {
      prec = xml_precision(read_record(f, r));
      after_format = true;
    }
    if (r.type == "scidac-binary-data" && after_format) {
      bin = &r;
      break;
    }
  }
  const qcd_uint_2 *L = v->geo->lL;
  const unsigned long V = v->geo->lV;
  const size_t w = prec == 64 ? 8 : 4;
  if ((prec != 64 && prec != 32) || !bin || bin->bytes < V * 24 * w) {
    std::fprintf(stderr, "process 0: Error in qcd_getVectorLime! Unsupported precision or missing/short binary data in %s\n",
                 fname);
    std::fclose(f);
    return 1;
  }
  std::vector<unsigned char> buf((size_t)V * 24 * w);
  std::fseek(f, bin->offset, SEEK_SET);
  const bool ok = std::fread(buf.data(), 1, buf.size(), f) == buf.size();
  std::fclose(f);
  if (!ok) {
    std::fprintf(stderr, "process 0: Error in qcd_getVectorLime! Short read of %s\n", fname);
    return 1;
  }
  const unsigned char *p = buf.data();
  for (int t = 0; t < L[0]; t++)
    for (int z = 0; z < L[3]; z++)
      for (int y = 0; y < L[2]; y++)
        for (int x = 0; x < L[1]; x++) {
          double *d = &v->D[lexic(t, x, y, z, L)][0][0].re;
          for (int k = 0; k < 24; k++, p += w) d[k] = w == 8 ? be_f64(p) : be_f32(p);
        }
  return 0;
}

int qcd_getVectorHMC(char *fname, qcd_uint_2 nu, qcd_uint_2 c2, qcd_vector *v) {
  if (!v || !v->initialized) {
    std::fprintf(stderr, "Error in qcd_getVectorHMC! Vector must be properly initialized.\n");
    return 1;
  }
  FILE *f = std::fopen(fname, "rb");
  if (!f) {
    std::fprintf(stderr, "process 0: Error in qcd_getVectorHMC! Cannot open %s for reading\n", fname);
    return 1;
  }
  double beta, kappa, mu;
  int fL, fT;
  if (std::fscanf(f, "%lf %lf %lf %d %d", &beta, &kappa, &mu, &fL, &fT) != 5 || std::fgetc(f) != '\n') {
    std::fprintf(stderr, "process 0: Error in qcd_getVectorHMC! Bad header line in %s\n", fname);
    std::fclose(f);
    return 1;
  }
  if (fL != v->geo->L[1] || fT != v->geo->L[0]) {
    std::fprintf(stderr, "process 0: Error in qcd_getVectorHMC! Dimensions in file %s don't match geometry!\n", fname);
    std::fclose(f);
    return 1;
  }
  const qcd_uint_2 *L = v->geo->lL;
  const unsigned long V = v->geo->lV;
  std::vector<unsigned char> buf((size_t)V * 24 * 8);
  const long start = std::ftell(f) + (long)((nu * 3 + c2) * V * 24 * 8);
  const bool ok = nu < 4 && c2 < 3 && std::fseek(f, start, SEEK_SET) == 0 &&
                  std::fread(buf.data(), 1, buf.size(), f) == buf.size();
  std::fclose(f);
  if (!ok) {
    std::fprintf(stderr, "process 0: Error in qcd_getVectorHMC! Short read of %s (nu %d, c2 %d)\n", fname, nu, c2);
    return 1;
  }
  const unsigned char *p = buf.data();
  for (int x = 0; x < L[1]; x++)
    for (int y = 0; y < L[2]; y++)
      for (int z = 0; z < L[3]; z++)
        for (int t = 0; t < L[0]; t++) {
          double *d = &v->D[lexic(t, x, y, z, L)][0][0].re;
          for (int k = 0; k < 24; k++, p += 8) d[k] = be_f64(p);
        }
  return 0;
}

int qcd_getVector(char *fname, int type, qcd_uint_2 nu, qcd_uint_2 c2, qcd_vector *v) {
  switch (type) {
    case qcd_PROP_HMC: return qcd_getVectorHMC(fname, nu, c2, v);
    case qcd_PROP_LIME:
    case qcd_SOURCE_LIME: return qcd_getVectorLime(fname, v);
    default:
      std::fprintf(stderr, "process 0: Error in qcd_getVector! Unknown type.\n");
      return 1;
  }
}

int qcd_writeVectorLime(char *fname, int type, qcd_vector *v) {
  if (!v || !v->initialized) {
    std::fprintf(stderr, "Error in qcd_writeVectorLime! Vector must be properly initialized.\n");
    return 1;
  }
  if (type != qcd_PROP_LIME && type != qcd_SOURCE_LIME) {
    std::fprintf(stderr, "Error in qcd_writeVectorLime! Unknown type.\n");
    return 1;
  }
  const qcd_uint_2 *L = v->geo->lL;
  char xml[512];
  std::snprintf(xml, sizeof(xml),
                "<?xml version=\"1.0\" encoding=\"UTF-8\"?>\n<etmcFormat>\n\t<field>diracFermion</field>\n"
                "\t<precision>64</precision>\n\t<flavours>1</flavours>\n\t<lx>%d</lx>\n\t<ly>%d</ly>\n\t<lz>%d</lz>\n"
                "\t<lt>%d</lt>\n\t<spin>4</spin>\n\t<colour>3</colour>\n</etmcFormat>",
                L[1], L[2], L[3], L[0]);
  const unsigned long V = v->geo->lV;
  std::vector<unsigned char> buf((size_t)V * 24 * 8);
  unsigned char *p = buf.data();
  for (int t = 0; t < L[0]; t++)
    for (int z = 0; z < L[3]; z++)
      for (int y = 0; y < L[2]; y++)
        for (int x = 0; x < L[1]; x++) {
          const double *d = &v->D[lexic(t, x, y, z, L)][0][0].re;
          for (int k = 0; k < 24; k++, p += 8) put_be_f64(p, d[k]);
        }
  FILE *f = std::fopen(fname, "wb");
  if (!f) {
    std::fprintf(stderr, "Error in qcd_writeVectorLime! Could not open %s for writing\n", fname);
    return 1;
  }
  const bool prop = type == qcd_PROP_LIME;
  const char *kind = prop ? "DiracFermion_Sink" : "DiracFermion_Source";
  LimeWriter w(f, 3);
  w.record(prop ? "propagator-type" : "source-type", kind, std::strlen(kind));
  w.record(prop ? "etmc-propagator-format" : "etmc-source-format", xml, std::strlen(xml));
  w.record("scidac-binary-data", buf.data(), buf.size());
  const bool ok = w.ok && std::fclose(f) == 0;
  if (!ok) std::fprintf(stderr, "Error in qcd_writeVectorLime! Write error on %s\n", fname);
  return ok ? 0 : 1;
}

int qcd_writeVector(char *fname, int type, qcd_uint_2 nu, qcd_uint_2 c2, qcd_vector *v) {
  (void)nu;
  (void)c2;
  if (!v || !v->initialized) {
    std::fprintf(stderr, "Error in qcd_writeVector! Vector must be properly initialized.\n");
    return 1;
  }
  if (type == qcd_PROP_LIME || type == qcd_SOURCE_LIME) return qcd_writeVectorLime(fname, type, v);
  std::fprintf(stderr, "process 0: Error in qcd_writeVector! Unknown type.\n");
  return 1;
}

// The HMC propagator file that qcd_getVectorHMC reads: a header line and the
// 12 vectors prop[3 nu + c2] (no reference writer exists for this format).
int qcd_writePropagatorHMC(char *fname, qcd_vector *vecs, qcd_real_8 beta, qcd_real_8 kappa, qcd_real_8 mu) {
  if (!vecs) return 1;
  for (int r = 0; r < 12; r++)
    if (!vecs[r].initialized) {
      std::fprintf(stderr, "Error in qcd_writePropagatorHMC! Vectors must be properly initialized.\n");
      return 1;
    }
  const qcd_uint_2 *L = vecs[0].geo->lL;
  const unsigned long V = vecs[0].geo->lV;
  FILE *f = std::fopen(fname, "wb");
  if (!f) {
    std::fprintf(stderr, "Error in qcd_writePropagatorHMC! Could not open %s for writing\n", fname);
    return 1;
  }
  bool ok = std::fprintf(f, "%.17g %.17g %.17g %d %d\n", beta, kappa, mu, L[1], L[0]) > 0;
  std::vector<unsigned char> buf((size_t)V * 24 * 8);
  for (int r = 0; r < 12 && ok; r++) {
    unsigned char *p = buf.data();
    for (int x = 0; x < L[1]; x++)
      for (int y = 0; y < L[2]; y++)
        for (int z = 0; z < L[3]; z++)
          for (int t = 0; t < L[0]; t++) {
            const double *d = &vecs[r].D[lexic(t, x, y, z, L)][0][0].re;
            for (int k = 0; k < 24; k++, p += 8) put_be_f64(p, d[k]);
          }
    ok = std::fwrite(buf.data(), 1, buf.size(), f) == buf.size();
  }
  ok = (std::fclose(f) == 0) && ok;
  if (!ok) std::fprintf(stderr, "Error in qcd_writePropagatorHMC! Write error on %s\n", fname);
  return ok ? 0 : 1;
}

}  // extern "C"
