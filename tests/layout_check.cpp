// layout_check.cpp -- host-side check of the cell-record layout of
// libminiamr_dev.so (make_record_layout, home, zcopy in amr_internal.h),
// built and run by tests/test_layout.py.
//
// For each shape, element size and one- or two-interior record it checks
//   * every canonical cell (i, j, k), 0 <= i <= nx+1 etc., has a home inside
//     the record, and no two cells share one;
//   * the z-plane copies of the cells at z = 1 and z = nz, and the second
//     interior and its z-plane copies, are inside the record and overlap no
//     home and no other copy;
//   * every part starts on a 128-byte line, the record is a multiple of
//     256 bytes, and a 16-cell row of the interior or of a z-plane copy
//     lies in one line when it is 128 bytes long.
// Prints one line per failure and exits 1 if there is any.
#include <cstdio>
#include <vector>

#include "amr_internal.h"

static int failures = 0;

#define CHECK(cond, ...)                                \
   do {                                                 \
      if (!(cond)) {                                    \
         if (failures++ < 20) {                         \
            std::printf("FAIL %s: ", #cond);            \
            std::printf(__VA_ARGS__);                   \
            std::printf("\n");                          \
         }                                              \
      }                                                 \
   } while (0)

static void check(int nx, int ny, int nz, int eb, bool two)
{
   const RecordLayout r = make_record_layout(nx, ny, nz, 27, eb, two);
   Lay L = r.g;
   const int64_t rec = L.rec;
   std::vector<int> owner(rec, 0);
   char tag[96];
   std::snprintf(tag, sizeof tag, "%dx%dx%d eb=%d two=%d", nx, ny, nz, eb, (int) two);
   auto claim = [&](int64_t off, const char *kind) {
      CHECK(off >= 0 && off < rec, "%s: %s offset %lld outside [0, %lld)", tag, kind,
            (long long) off, (long long) rec);
      if (off >= 0 && off < rec) {
         CHECK(owner[off] == 0, "%s: %s offset %lld already taken", tag, kind, (long long) off);
         owner[off]++;
      }
   };
   const int copies = two ? 2 : 1;
   for (int c = 0; c < copies; c++) {
      L.oI = r.oI_par[c];
      L.oZI = r.oZI_par[c];
      for (int i = 0; i <= nx + 1; i++)
         for (int j = 0; j <= ny + 1; j++)
            for (int k = 0; k <= nz + 1; k++) {
               const bool inner = i >= 1 && i <= nx && j >= 1 && j <= ny && k >= 1 && k <= nz;
               // ghost homes and ghost z copies are single: claim them once
               if (inner || c == 0)
                  claim(home(L, i, j, k), "home");
               if ((k == 1 || k == nz) && (inner || c == 0))
                  claim(zcopy(L, i, j, k), "zcopy");
            }
   }
   const int line = 128 / eb;
   const int parts[] = {L.oF[0], L.oF[1], L.oF[2], L.oF[3], L.oF[4], L.oF[5], L.oEXY,
                        L.oEXZ,  L.oEYZ,  L.oZX,   L.oZY,   r.oI_par[0], r.oI_par[1],
                        r.oZI_par[0], r.oZI_par[1]};
   for (int p : parts)
      CHECK(p % line == 0, "%s: part at %d not on a 128-byte line", tag, p);
   CHECK((rec * eb) % 256 == 0, "%s: record of %lld elements not a multiple of 256 bytes", tag,
         (long long) rec);
   if (nz * eb == 128)
      for (int c = 0; c < copies; c++)
         CHECK((r.oI_par[c] * eb) % 128 == 0, "%s: interior rows straddle lines", tag);
   if (ny * eb == 128)
      for (int c = 0; c < copies; c++)
         CHECK((r.oZI_par[c] * eb) % 128 == 0, "%s: z-copy rows straddle lines", tag);
}

int main()
{
   const int shapes[][3] = {{16, 16, 16}, {8, 8, 8}, {4, 6, 10}, {12, 8, 16}, {2, 2, 2},
                            {6, 4, 2}, {32, 16, 8}, {10, 12, 14}};
   int cases = 0;
   for (auto &s : shapes)
      for (int eb : {8, 4, 2})
         for (bool two : {false, true}) {
            check(s[0], s[1], s[2], eb, two);
            cases++;
         }
   std::printf("%d layouts checked, %d failures\n", cases, failures);
   return failures ? 1 : 0;
}
