// codec.hip -- the halo message codec: pack a block face into a message
// buffer and unpack a received message into a block's ghost plane, in the
// reference's message formats (pack_face / unpack_face, ref/comm.c:245-990
// and :993-1461; lengths ref/init.c:77-117).  On one GPU every face is
// exchanged in place (exchange.hip); these kernels are the one-GPU part of
// a multi-GPU exchange, where the messages cross devices (RCCL send/recv of
// the packed buffers, one message per off-device face).
//
// Formats as data.  The reference writes every (code, direction, face
// case) as its own loop nest.  Each of them walks one face plane P[a][b]
// (a, b = the two in-plane axes in the reference's loop order: j, k for x
// faces; i, k for y; i, j for z) as a product of two axis sequences: the
// message element (ea, eb) covers cells A[ea] x B[eb], where each entry of
// A and B is one coordinate or a pair (c, c+1).  pack sums those cells in
// the reference's term order (a outer, b inner: [a][b] + [a][b+1] +
// [a+1][b] + [a+1][b+1]) and divides by 4.0 where the reference does;
// unpack writes the element's value to every cell it covers.  The host
// builds the sequences of all 3 codes x 3 directions x 20 face cases x
// {pack, unpack} once per context (build_formats), and one kernel per side
// serves them all.  A message holds its num_comm variables one after
// another, each |A| x |B| values in row order, in the storage type.
#include <algorithm>
#include <cstring>

#include "amr_internal.h"

namespace {

struct CodecFmt {
   int32_t plane;   // index of the plane along the direction (0..n+1)
   int32_t div4;    // pack divides each element by 4.0
   int32_t na, nb;  // entries of the two axis sequences
   int32_t a_off, b_off;   // their first entry in the group table
};

constexpr int kFormats = 3 * 3 * 20 * 2;

__host__ __device__ __forceinline__ int fmt_index(int code, int dir, int face_case, int side)
{
   return ((code * 3 + dir) * 20 + face_case) * 2 + side;
}

// Group entry: coordinate c in the low 16 bits, bit 16 set for the pair
// (c, c+1).
struct Seq {
   std::vector<int32_t> g;
   void one(int c) { g.push_back(c); }
   void pair(int c) { g.push_back(c | (1 << 16)); }
   void range(int s, int e) { for (int c = s; c <= e; c++) one(c); }
   void pairs(int n) { for (int c = 1; c < n; c += 2) pair(c); }
};

// The face plane's extents along (a, b) and their halves.
struct Ext {
   int na, nb, ha, hb, nd;
};

Ext extents(const Lay &L, int dir)
{
   const int n[3] = { L.nx, L.ny, L.nz };
   Ext e;
   e.na = dir == 0 ? L.ny : L.nx;
   e.nb = dir == 2 ? L.ny : L.nz;
   e.ha = e.na / 2;
   e.hb = e.nb / 2;
   e.nd = n[dir];
   return e;
}

// pack_face, ref/comm.c:254-403 (code 0), :403-657 (code 1), :657-990
// (code 2); c is the face case without the +10 of E/N/U faces.
void pack_format(int code, int dir, int c, const Ext &x, Seq &A, Seq &B, int &div4)
{
   div4 = 0;
   const bool lo_a = c % 2 == 0, lo_b = (c / 2) % 2 == 1;
   if (code == 0) {
      if (c < 2) {                 // x: interior for both; y/z case 1 add the ring
         const bool ra = c == 1 && dir > 0, rb = c == 1 && dir == 2;
         A.range(ra ? 0 : 1, ra ? x.na + 1 : x.na);
         B.range(rb ? 0 : 1, rb ? x.nb + 1 : x.nb);
      } else if (c <= 5) {         // fine whole face -> coarse quarter: 2x2 sums
         A.pairs(x.na);
         B.pairs(x.nb);
      } else {                     // coarse quarter -> fine whole face: / 4
         A.range(lo_a ? 1 : x.ha + 1, lo_a ? x.ha : x.na);
         B.range(lo_b ? 1 : x.hb + 1, lo_b ? x.hb : x.nb);
         div4 = 1;
      }
      return;
   }
   if (c < 2) {                    // whole ghost planes
      A.range(0, x.na + 1);
      B.range(0, x.nb + 1);
   } else if (c <= 5) {            // 2x2 sums plus this quarter's share of the ring
      if (lo_a) A.one(0);
      A.pairs(x.na);
      if (!lo_a) A.one(x.na + 1);
      if (lo_b) B.one(0);
      B.pairs(x.nb);
      if (!lo_b) B.one(x.nb + 1);
   } else if (code == 1) {         // the quarter with its ring row and column, / 4
      A.range(lo_a ? 0 : x.ha, lo_a ? x.ha + 1 : x.na + 1);
      B.range(lo_b ? 0 : x.hb, lo_b ? x.hb + 1 : x.nb + 1);
      div4 = 1;
   } else {                        // code 2: prolonged on the send side, / 4
      const int a0 = lo_a ? 1 : x.ha + 1, a1 = lo_a ? x.ha : x.na;
      const int b0 = lo_b ? 1 : x.hb + 1, b1 = lo_b ? x.hb : x.nb;
      A.one(a0 - 1);
      for (int a = a0; a <= a1; a++) { A.one(a); A.one(a); }
      A.one(a1 + 1);
      B.one(b0 - 1);
      for (int b = b0; b <= b1; b++) { B.one(b); B.one(b); }
      B.one(b1 + 1);
      div4 = 1;
   }
}

// unpack_face, ref/comm.c:1002-1152 (code 0), :1152-1356 (code 1),
// :1356-1461 (code 2).
void unpack_format(int code, int dir, int c, const Ext &x, Seq &A, Seq &B)
{
   const bool lo_a = c % 2 == 0, lo_b = (c / 2) % 2 == 1;
   if (code == 0) {
      if (c < 2) {
         const bool ra = c == 1 && dir > 0, rb = c == 1 && dir == 2;
         A.range(ra ? 0 : 1, ra ? x.na + 1 : x.na);
         B.range(rb ? 0 : 1, rb ? x.nb + 1 : x.nb);
      } else if (c <= 5) {         // fine side: each value into 2x2 cells
         A.pairs(x.na);
         B.pairs(x.nb);
      } else {                     // coarse side: one quarter
         A.range(lo_a ? 1 : x.ha + 1, lo_a ? x.ha : x.na);
         B.range(lo_b ? 1 : x.hb + 1, lo_b ? x.hb : x.nb);
      }
      return;
   }
   if (c < 2 || (code == 2 && c <= 5)) {
      A.range(0, x.na + 1);
      B.range(0, x.nb + 1);
   } else if (c <= 5) {            // code 1 fine side: ring cells once, the rest 2x2
      A.one(0);
      A.pairs(x.na);
      A.one(x.na + 1);
      B.one(0);
      B.pairs(x.nb);
      B.one(x.nb + 1);
   } else {                        // coarse side: the quarter and its ring share
      A.range(lo_a ? 0 : x.ha + 1, lo_a ? x.ha : x.na + 1);
      B.range(lo_b ? 0 : x.hb + 1, lo_b ? x.hb : x.nb + 1);
   }
}

int build_formats(amr_dev *d)
{
   if (d->d_codec)
      return AMR_OK;
   std::vector<CodecFmt> f(kFormats);
   std::vector<int32_t> groups;
   for (int code = 0; code < 3; code++)
      for (int dir = 0; dir < 3; dir++) {
         const Ext x = extents(d->g, dir);
         for (int fc = 0; fc < 20; fc++)
            for (int side = 0; side < 2; side++) {
               const int c = fc % 10;
               Seq A, B;
               int div4 = 0;
               if (side == 0)
                  pack_format(code, dir, c, x, A, B, div4);
               else
                  unpack_format(code, dir, c, x, A, B);
               CodecFmt &t = f[fmt_index(code, dir, fc, side)];
               // pack reads plane 1 / n; unpack writes ghost plane 0 / n+1
               t.plane = side == 0 ? (fc >= 10 ? x.nd : 1) : (fc >= 10 ? x.nd + 1 : 0);
               t.div4 = div4;
               t.na = (int) A.g.size();
               t.nb = (int) B.g.size();
               t.a_off = (int) groups.size();
               groups.insert(groups.end(), A.g.begin(), A.g.end());
               t.b_off = (int) groups.size();
               groups.insert(groups.end(), B.g.begin(), B.g.end());
            }
      }
   const size_t bytes = sizeof(CodecFmt) * f.size() + sizeof(int32_t) * groups.size();
   std::vector<char> host(bytes);
   std::memcpy(host.data(), f.data(), sizeof(CodecFmt) * f.size());
   std::memcpy(host.data() + sizeof(CodecFmt) * f.size(), groups.data(),
          sizeof(int32_t) * groups.size());
   for (size_t i = 0; i < f.size(); i++)
      d->codec_len[i] = f[i].na * f[i].nb;
   HIP_TRY(d, hipMalloc(&d->d_codec, bytes));
   HIP_TRY(d, hipMemcpy(d->d_codec, host.data(), bytes, hipMemcpyHostToDevice));
   return AMR_OK;
}

// Canonical cell of plane coordinate (a, b) at index p along dir.
__device__ __forceinline__ void cell_of(int dir, int p, int a, int b, int &i, int &j, int &k)
{
   i = dir == 0 ? p : a;
   j = dir == 0 ? a : dir == 1 ? p : b;
   k = dir == 2 ? p : b;
}

// One workgroup per (message, variable).
template <typename T>
__global__ __launch_bounds__(128) void k_pack(const T *__restrict__ pool, Lay L,
                                              const CodecFmt *__restrict__ fmts,
                                              const amr_face_msg *__restrict__ msgs, int base,
                                              int dir, int start, T *__restrict__ buf)
{
   typedef typename Acc<T>::type A;
   const amr_face_msg m = msgs[blockIdx.x];
   const CodecFmt f = fmts[base + m.face_case * 2];
   const int32_t *groups = (const int32_t *) (fmts + kFormats);
   const int len = f.na * f.nb;
   const T *rec = pool + (int64_t) (start + blockIdx.y) * L.var_stride + (int64_t) m.slot * L.rec;
   T *out = buf + m.offset + (int64_t) blockIdx.y * len;
   for (int e = threadIdx.x; e < len; e += blockDim.x) {
      const int ea = e / f.nb, eb = e - ea * f.nb;
      const int ga = groups[f.a_off + ea], gb = groups[f.b_off + eb];
      const int a0 = ga & 0xffff, a1 = a0 + (ga >> 16), b0 = gb & 0xffff, b1 = b0 + (gb >> 16);
      A s = A(0);
      bool first = true;
      for (int a = a0; a <= a1; a++)
         for (int b = b0; b <= b1; b++) {
            int i, j, k;
            cell_of(dir, f.plane, a, b, i, j, k);
            const A v = to_acc(rec[home(L, i, j, k)]);
            s = first ? v : s + v;
            first = false;
         }
      if (f.div4)
         s = s / A(4);
      out[e] = from_acc<T>(s);
   }
}

template <typename T>
__global__ __launch_bounds__(128) void k_unpack(T *__restrict__ pool, Lay L,
                                                const CodecFmt *__restrict__ fmts,
                                                const amr_face_msg *__restrict__ msgs, int base,
                                                int dir, int start, const T *__restrict__ buf)
{
   const amr_face_msg m = msgs[blockIdx.x];
   const CodecFmt f = fmts[base + m.face_case * 2 + 1];
   const int32_t *groups = (const int32_t *) (fmts + kFormats);
   const int len = f.na * f.nb;
   T *rec = pool + (int64_t) (start + blockIdx.y) * L.var_stride + (int64_t) m.slot * L.rec;
   const T *in = buf + m.offset + (int64_t) blockIdx.y * len;
   for (int e = threadIdx.x; e < len; e += blockDim.x) {
      const int ea = e / f.nb, eb = e - ea * f.nb;
      const int ga = groups[f.a_off + ea], gb = groups[f.b_off + eb];
      const int a0 = ga & 0xffff, a1 = a0 + (ga >> 16), b0 = gb & 0xffff, b1 = b0 + (gb >> 16);
      const T v = in[e];
      for (int a = a0; a <= a1; a++)
         for (int b = b0; b <= b1; b++) {
            int i, j, k;
            cell_of(dir, f.plane, a, b, i, j, k);
            store_cell(rec, L, i, j, k, v);
         }
   }
}

template <typename T>
int launch(amr_dev *d, bool pack, const Lay &L, int base, int dir, int start, int num, int n,
           void *buf)
{
   const CodecFmt *fmts = (const CodecFmt *) d->d_codec;
   const amr_face_msg *msgs = (const amr_face_msg *) d->d_msgs;
   dim3 grid(n, num);
   if (pack)
      hipLaunchKernelGGL(k_pack<T>, grid, dim3(128), 0, d->stream, (const T *) d->pool, L, fmts,
                         msgs, base, dir, start, (T *) buf);
   else
      hipLaunchKernelGGL(k_unpack<T>, grid, dim3(128), 0, d->stream, (T *) d->pool, L, fmts,
                         msgs, base, dir, start, (const T *) buf);
   HIP_TRY(d, hipGetLastError());
   return AMR_OK;
}

int codec(amr_dev *d, bool pack, int code, int dir, int start, int num,
          const amr_face_msg *msgs, int n, void *buf, size_t buf_elems)
{
   const char *what = pack ? "pack_faces" : "unpack_faces";
   if (!d)
      return AMR_ERR_ARG;
   if (code < 0 || code > 2 || dir < 0 || dir > 2 || start < 0 || num < 1 ||
       start + num > d->c.num_vars || n < 0 || (n && (!msgs || !buf)))
      return amr_set_error(d, AMR_ERR_ARG, "%s: bad arguments", what);
   if (!n)
      return AMR_OK;
   if (amr_require_fresh(d, start, num, what))
      return AMR_ERR_STATE;
   if (int rc = build_formats(d))
      return rc;
   for (int t = 0; t < n; t++) {
      const amr_face_msg &m = msgs[t];
      if (m.slot < 0 || m.slot >= d->c.max_blocks || m.face_case < 0 || m.face_case > 19 ||
          m.offset < 0)
         return amr_set_error(d, AMR_ERR_ARG, "%s: bad message %d", what, t);
      const int len = d->codec_len[fmt_index(code, dir, m.face_case, pack ? 0 : 1)];
      if ((size_t) m.offset + (size_t) num * len > buf_elems)
         return amr_set_error(d, AMR_ERR_ARG, "%s: message %d ends past the buffer (%zu values)",
                              what, t, buf_elems);
   }
   if ((size_t) n * sizeof(amr_face_msg) > d->msgs_bytes) {
      HIP_TRY(d, hipStreamSynchronize(d->stream));
      hipFree(d->d_msgs);
      d->d_msgs = nullptr;
      d->msgs_bytes = 0;
      HIP_TRY(d, hipMalloc(&d->d_msgs, (size_t) n * sizeof(amr_face_msg)));
      d->msgs_bytes = (size_t) n * sizeof(amr_face_msg);
   }
   HIP_TRY(d, hipStreamSynchronize(d->stream));   // earlier calls may still read d_msgs
   HIP_TRY(d, hipMemcpy(d->d_msgs, msgs, (size_t) n * sizeof(amr_face_msg),
                        hipMemcpyHostToDevice));
   Lay L;
   if (int rc = amr_lay(d, start, num, &L))
      return rc;
   const int base = fmt_index(code, dir, 0, 0);
   amr_span_begin(d, 0);
   int rc;
   switch (d->c.precision) {
   case AMR_PREC_F64: rc = launch<double>(d, pack, L, base, dir, start, num, n, buf); break;
   case AMR_PREC_F32: rc = launch<float>(d, pack, L, base, dir, start, num, n, buf); break;
   default: rc = launch<uint16_t>(d, pack, L, base, dir, start, num, n, buf); break;
   }
   amr_span_end(d);
   return rc;
}

}  // namespace

extern "C" {

int amr_dev_face_msg_len(amr_dev *d, int code, int dir, int face_case, int unpack)
{
   if (!d)
      return AMR_ERR_ARG;
   if (code < 0 || code > 2 || dir < 0 || dir > 2 || face_case < 0 || face_case > 19)
      return amr_set_error(d, AMR_ERR_ARG, "face_msg_len: bad arguments");
   if (int rc = build_formats(d))
      return rc;
   return d->codec_len[fmt_index(code, dir, face_case, unpack ? 1 : 0)];
}

int amr_dev_pack_faces(amr_dev *d, int code, int dir, int start, int num_comm,
                       const amr_face_msg *msgs, int n, void *buf, size_t buf_elems)
{
   return codec(d, true, code, dir, start, num_comm, msgs, n, buf, buf_elems);
}

int amr_dev_unpack_faces(amr_dev *d, int code, int dir, int start, int num_comm,
                         const amr_face_msg *msgs, int n, const void *buf, size_t buf_elems)
{
   return codec(d, false, code, dir, start, num_comm, msgs, n, const_cast<void *>(buf),
                buf_elems);
}

}  // extern "C"
