// fgpe_ckpt.hip -- binary checkpoint/resume of a run (fgpe_checkpoint_*).
//
// The reference can only restart real time, from the text file initial_sol
// (SRC/cgpe3d.f90:383-410), which keeps 10 decimals of density and phase and
// loses the loop state.  A checkpoint here keeps what a bit-identical
// continuation needs: the wave function as the device holds it (FP32 or FP64
// complex, [m][z][y][x]), the device scalars carried from step to step (the
// NORMC of the last step and the spin-2 projection's Newton state), and the
// caller's loop state (II, EN_OLD, six free values).  The run parameters are
// stored too, and a load into a handle whose parameters differ is refused.
//
// File layout, little-endian, fixed sizes, no padding:
//   char[8]  "FGPECKP1"
//   int32    version (1), dim, spin, precision, imag, soc, n[3]
//   double   d[3], dt, gamma[3], tau[3], alpha[3], mag
//   int64    ii;  double en_old, user[6]
//   double   norm[8], sig[2]                  (device scalars)
//   int64    payload bytes (NX NY NZ MS * 8 or 16)
//   payload
//   uint64   FNV-1a 64 of every byte before it
#include "fgpe_impl.hpp"

using namespace fgpe_impl;

namespace {

constexpr char MAGIC[8] = {'F', 'G', 'P', 'E', 'C', 'K', 'P', '1'};
constexpr int32_t VERSION = 1;

struct Fnv {
  uint64_t h = 0xcbf29ce484222325ULL;
  void add(const void* p, size_t n) {
    const unsigned char* b = (const unsigned char*)p;
    for (size_t i = 0; i < n; ++i) {
      h ^= b[i];
      h *= 0x100000001b3ULL;
    }
  }
};

// Header fields in file order (everything before the payload).
struct Header {
  int32_t version, dim, spin, precision, imag, soc, n[3];
  double d[3], dt, gamma[3], tau[3], alpha[3], mag;
  int64_t ii;
  double en_old, user[6];
  double norm[8], sig[2];
  int64_t payload;
};

template <typename V> void put(std::vector<unsigned char>& b, const V& v) {
  const unsigned char* p = (const unsigned char*)&v;
  b.insert(b.end(), p, p + sizeof(V));
}
template <typename V> bool get(const unsigned char*& p, const unsigned char* end, V& v) {
  if ((size_t)(end - p) < sizeof(V)) return false;
  std::memcpy(&v, p, sizeof(V));
  p += sizeof(V);
  return true;
}

std::vector<unsigned char> encode(const Header& h) {
  std::vector<unsigned char> b(MAGIC, MAGIC + 8);
  put(b, h.version);
  put(b, h.dim);
  put(b, h.spin);
  put(b, h.precision);
  put(b, h.imag);
  put(b, h.soc);
  for (int i = 0; i < 3; ++i) put(b, h.n[i]);
  for (int i = 0; i < 3; ++i) put(b, h.d[i]);
  put(b, h.dt);
  for (int i = 0; i < 3; ++i) put(b, h.gamma[i]);
  for (int i = 0; i < 3; ++i) put(b, h.tau[i]);
  for (int i = 0; i < 3; ++i) put(b, h.alpha[i]);
  put(b, h.mag);
  put(b, h.ii);
  put(b, h.en_old);
  for (int i = 0; i < 6; ++i) put(b, h.user[i]);
  for (int i = 0; i < 8; ++i) put(b, h.norm[i]);
  for (int i = 0; i < 2; ++i) put(b, h.sig[i]);
  put(b, h.payload);
  return b;
}

bool decode(const std::vector<unsigned char>& b, Header& h) {
  if (b.size() < 8 || std::memcmp(b.data(), MAGIC, 8) != 0) return false;
  const unsigned char* p = b.data() + 8;
  const unsigned char* e = b.data() + b.size();
  bool ok = get(p, e, h.version) && get(p, e, h.dim) && get(p, e, h.spin) && get(p, e, h.precision) &&
            get(p, e, h.imag) && get(p, e, h.soc);
  for (int i = 0; i < 3; ++i) ok = ok && get(p, e, h.n[i]);
  for (int i = 0; i < 3; ++i) ok = ok && get(p, e, h.d[i]);
  ok = ok && get(p, e, h.dt);
  for (int i = 0; i < 3; ++i) ok = ok && get(p, e, h.gamma[i]);
  for (int i = 0; i < 3; ++i) ok = ok && get(p, e, h.tau[i]);
  for (int i = 0; i < 3; ++i) ok = ok && get(p, e, h.alpha[i]);
  ok = ok && get(p, e, h.mag) && get(p, e, h.ii) && get(p, e, h.en_old);
  for (int i = 0; i < 6; ++i) ok = ok && get(p, e, h.user[i]);
  for (int i = 0; i < 8; ++i) ok = ok && get(p, e, h.norm[i]);
  for (int i = 0; i < 2; ++i) ok = ok && get(p, e, h.sig[i]);
  return ok && get(p, e, h.payload) && p == e;
}

size_t header_bytes() { return encode(Header{}).size(); }

Header from_params(const fgpe_params& p) {
  Header h{};
  h.version = VERSION;
  h.dim = p.dim;
  h.spin = p.spin;
  h.precision = p.precision;
  h.imag = p.imag;
  h.soc = p.soc;
  for (int i = 0; i < 3; ++i) {
    h.n[i] = p.n[i];
    h.d[i] = p.d[i];
    h.gamma[i] = p.gamma[i];
    h.tau[i] = p.tau[i];
    h.alpha[i] = p.alpha[i];
  }
  h.dt = p.dt;
  h.mag = p.mag;
  return h;
}

struct File {
  FILE* f = nullptr;
  explicit File(FILE* x) : f(x) {}
  ~File() {
    if (f) fclose(f);
  }
};

// Reads the header and verifies the checksum of the whole file, streaming the
// payload in blocks; with dst != nullptr the payload is also copied there.
int read_checked(const char* path, Header& h, unsigned char* dst, size_t dst_bytes) {
  File in(fopen(path, "rb"));
  if (!in.f) return fail(FGPE_EINVAL, "cannot open checkpoint '%s'", path);
  const size_t hb = header_bytes();
  std::vector<unsigned char> head(hb);
  if (fread(head.data(), 1, hb, in.f) != hb || !decode(head, h))
    return fail(FGPE_EINVAL, "'%s' is not a libfgpe checkpoint", path);
  if (h.version != VERSION) return fail(FGPE_EINVAL, "checkpoint version %d (expected %d)", h.version, VERSION);
  if (h.payload < 0 || (dst && (size_t)h.payload != dst_bytes))
    return fail(FGPE_EINVAL, "checkpoint payload of %lld bytes, expected %zu", (long long)h.payload, dst_bytes);
  Fnv sum;
  sum.add(head.data(), hb);
  std::vector<unsigned char> buf(1 << 22);
  int64_t left = h.payload;
  unsigned char* out = dst;
  while (left > 0) {
    const size_t k = (size_t)std::min<int64_t>(left, (int64_t)buf.size());
    unsigned char* tgt = out ? out : buf.data();
    if (fread(tgt, 1, k, in.f) != k) return fail(FGPE_EINVAL, "checkpoint '%s' is truncated", path);
    sum.add(tgt, k);
    if (out) out += k;
    left -= (int64_t)k;
  }
  uint64_t stored = 0;
  if (fread(&stored, 1, sizeof stored, in.f) != sizeof stored)
    return fail(FGPE_EINVAL, "checkpoint '%s' is truncated", path);
  if (fgetc(in.f) != EOF) return fail(FGPE_EINVAL, "checkpoint '%s' has trailing bytes", path);
  if (stored != sum.h) return fail(FGPE_EINVAL, "checkpoint '%s' fails its checksum", path);
  return 0;
}

}  // namespace

extern "C" int fgpe_checkpoint_save(fgpe_handle* h, const char* path, const fgpe_loop_state* st) {
  if (!h || !path || !st) return fail(FGPE_EINVAL, "null argument");
  Header hd = from_params(h->p);
  hd.ii = st->ii;
  hd.en_old = st->en_old;
  for (int i = 0; i < 6; ++i) hd.user[i] = st->user[i];
  const size_t bytes = (size_t)h->npts * h->ms * (h->fp32 ? sizeof(float2) : sizeof(double2));
  hd.payload = (int64_t)bytes;
  std::vector<unsigned char> data(bytes);
  HIPCHK(hipStreamSynchronize(h->stream));
  HIPCHK(hipMemcpy(hd.norm, h->dnorm, sizeof hd.norm, hipMemcpyDeviceToHost));
  HIPCHK(hipMemcpy(hd.sig, h->dsig, sizeof hd.sig, hipMemcpyDeviceToHost));
  HIPCHK(hipMemcpy(data.data(), h->phi, bytes, hipMemcpyDeviceToHost));
  const std::vector<unsigned char> head = encode(hd);
  Fnv sum;
  sum.add(head.data(), head.size());
  sum.add(data.data(), data.size());
  // write to a temporary name, then rename: a crash never leaves half a file
  const std::string tmp = std::string(path) + ".tmp";
  {
    File out(fopen(tmp.c_str(), "wb"));
    if (!out.f) return fail(FGPE_EINVAL, "cannot write checkpoint '%s'", tmp.c_str());
    if (fwrite(head.data(), 1, head.size(), out.f) != head.size() ||
        fwrite(data.data(), 1, data.size(), out.f) != data.size() ||
        fwrite(&sum.h, 1, sizeof sum.h, out.f) != sizeof sum.h || fflush(out.f) != 0)
      return fail(FGPE_EINVAL, "short write of checkpoint '%s'", tmp.c_str());
  }
  if (std::rename(tmp.c_str(), path) != 0) return fail(FGPE_EINVAL, "cannot rename '%s' to '%s'", tmp.c_str(), path);
  return 0;
}

extern "C" int fgpe_checkpoint_load(fgpe_handle* h, const char* path, fgpe_loop_state* st) {
  if (!h || !path || !st) return fail(FGPE_EINVAL, "null argument");
  const size_t bytes = (size_t)h->npts * h->ms * (h->fp32 ? sizeof(float2) : sizeof(double2));
  Header hd{};
  std::vector<unsigned char> data(bytes);
  int rc = read_checked(path, hd, data.data(), bytes);
  if (rc) return rc;
  const Header want = from_params(h->p);
  if (hd.dim != want.dim || hd.spin != want.spin || hd.precision != want.precision || hd.imag != want.imag ||
      hd.soc != want.soc)
    return fail(FGPE_EINVAL, "checkpoint of a dim %d spin %d precision %d imag %d run, handle is %d/%d/%d/%d", hd.dim,
                hd.spin, hd.precision, hd.imag, want.dim, want.spin, want.precision, want.imag);
  for (int i = 0; i < 3; ++i)
    if (hd.n[i] != want.n[i]) return fail(FGPE_EINVAL, "checkpoint grid differs on axis %d", i + 1);
  // the physics must be the same to the bit: a resumed run is the same run
  auto same = [](double a, double b) { return std::memcmp(&a, &b, sizeof a) == 0; };
  bool ok = same(hd.dt, want.dt) && same(hd.mag, want.mag);
  for (int i = 0; i < 3; ++i)
    ok = ok && same(hd.d[i], want.d[i]) && same(hd.gamma[i], want.gamma[i]) && same(hd.tau[i], want.tau[i]) &&
         same(hd.alpha[i], want.alpha[i]);
  if (!ok) return fail(FGPE_EINVAL, "checkpoint parameters (DT, DX, GAMMA, TAU, ALPHA or MAG) differ from the handle's");
  HIPCHK(hipStreamSynchronize(h->stream));
  HIPCHK(hipMemcpy(h->phi, data.data(), bytes, hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(h->dnorm, hd.norm, sizeof hd.norm, hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(h->dsig, hd.sig, sizeof hd.sig, hipMemcpyHostToDevice));
  st->ii = hd.ii;
  st->en_old = hd.en_old;
  for (int i = 0; i < 6; ++i) st->user[i] = hd.user[i];
  return 0;
}

extern "C" int fgpe_checkpoint_info(const char* path, fgpe_checkpoint_header* info) {
  if (!path || !info) return fail(FGPE_EINVAL, "null argument");
  Header hd{};
  int rc = read_checked(path, hd, nullptr, 0);
  if (rc) return rc;
  info->version = hd.version;
  info->dim = hd.dim;
  info->spin = hd.spin;
  info->precision = hd.precision;
  info->imag = hd.imag;
  for (int i = 0; i < 3; ++i) info->n[i] = hd.n[i];
  info->dt = hd.dt;
  info->ii = hd.ii;
  info->en_old = hd.en_old;
  info->payload_bytes = hd.payload;
  return 0;
}
