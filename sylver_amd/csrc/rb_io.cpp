// Rutherford-Boeing reader for the spldlt_test driver.
//
// The reference's drivers read their matrices with SPRAL's rb_read
// (drivers/spldlt_test.F90:86-95, rb_options%values = 2: symmetric values
// made up when the file holds a pattern only); SPRAL is not vendored, so this
// is a reader of the published Rutherford-Boeing / Harwell-Boeing layout:
//
//   line 1  title (72 columns) and key (8)
//   line 2  card counts: total, pointer, index, value [, right-hand side]
//   line 3  matrix type (3 letters: r/p/i, s/u/h/z/r, a/e), nrow, ncol,
//           nnz, neltvl
//   line 4  Fortran formats of the pointers, indices, values [, rhs]
//   line 5  (Harwell-Boeing, rhs cards > 0 only) right-hand side header
//   then the column pointers, row indices and values in those fixed-width
//   formats (1-based).
//
// Output: the lower triangle in 1-based CSC (the layout spldlt_analyse
// takes), int64 pointers. Symmetric types keep their entries (an entry above
// the diagonal is moved to its transposed position); unsymmetric and
// rectangular assembled types keep the entries with row >= column of a square
// matrix. Complex and elemental matrices are not supported.
#include <algorithm>
#include <cctype>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/sylver_amd.h"

namespace {

struct Field {
   int per_line = 0;  // items per line
   int width = 0;     // characters per item
};

// "(13I6)", "(1P,4E20.12)", "(1P4D25.16)", "(5E16.8)", "(8F10.3)"
bool parse_format(const std::string& f, Field& out) {
   std::string s;
   for (char c : f)
      if (!std::isspace((unsigned char)c)) s += (char)std::toupper((unsigned char)c);
   size_t a = s.find('('), b = s.rfind(')');
   if (a == std::string::npos || b == std::string::npos || b <= a) return false;
   s = s.substr(a + 1, b - a - 1);
   // drop scale factors "kP" (alone or glued to the descriptor)
   size_t i = 0;
   while (i < s.size()) {
      size_t j = i;
      while (j < s.size() && std::isdigit((unsigned char)s[j])) ++j;
      if (j < s.size() && s[j] == 'P') {
         i = j + 1;
         if (i < s.size() && s[i] == ',') ++i;
         continue;
      }
      break;
   }
   s = s.substr(i);
   size_t p = 0;
   int rep = 0;
   while (p < s.size() && std::isdigit((unsigned char)s[p])) rep = rep * 10 + (s[p++] - '0');
   if (rep == 0) rep = 1;
   if (p >= s.size() || std::string("IEDFG").find(s[p]) == std::string::npos) return false;
   ++p;
   int w = 0;
   while (p < s.size() && std::isdigit((unsigned char)s[p])) w = w * 10 + (s[p++] - '0');
   if (w <= 0) return false;
   out.per_line = rep;
   out.width = w;
   return true;
}

// Fortran real field: D exponents, and the exponent letter may be missing
// ("1.25-3" is 1.25e-3).
bool parse_real(const char* p, int w, double& v) {
   char buf[96];
   int k = 0;
   for (int i = 0; i < w && p[i] && k < 90; ++i) {
      char c = p[i];
      if (c == '\n' || c == '\r') break;
      if (c == ' ') continue;
      if (c == 'D' || c == 'd' || c == 'Q' || c == 'q') c = 'E';
      if ((c == '+' || c == '-') && k > 0 && buf[k - 1] != 'E' && buf[k - 1] != 'e') buf[k++] = 'E';
      buf[k++] = c;
   }
   if (k == 0) return false;
   buf[k] = 0;
   char* end = nullptr;
   v = std::strtod(buf, &end);
   return end && *end == 0;
}

bool parse_int(const char* p, int w, int64_t& v) {
   char buf[40];
   int k = 0;
   for (int i = 0; i < w && p[i] && k < 38; ++i) {
      char c = p[i];
      if (c == '\n' || c == '\r') break;
      if (c == ' ') continue;
      buf[k++] = c;
   }
   if (k == 0) return false;
   buf[k] = 0;
   char* end = nullptr;
   v = std::strtoll(buf, &end, 10);
   return end && *end == 0;
}

struct Lines {
   FILE* f;
   std::string line;
   bool next() {
      line.clear();
      int c;
      bool any = false;
      while ((c = std::fgetc(f)) != EOF) {
         any = true;
         if (c == '\n') break;
         line += (char)c;
      }
      if (!line.empty() && line.back() == '\r') line.pop_back();
      return any;
   }
};

// One line holds min(per_line, remaining) items. Fields are fixed-width; a
// line whose blank-separated tokens are exactly that many items is read token
// by token instead, which also accepts writers that pad a field one column
// short of its declared width (a packed field, e.g. "10001001" for two I4
// items, never splits into the right number of tokens).
template <typename T, typename Parse>
bool read_block(Lines& in, const Field& fmt, int64_t count, std::vector<T>& out, Parse parse) {
   out.resize(count);
   int64_t k = 0;
   while (k < count) {
      if (!in.next()) return false;
      const std::string& l = in.line;
      const int64_t want = std::min<int64_t>(fmt.per_line, count - k);
      std::vector<std::pair<size_t, size_t>> tok;
      for (size_t p = 0; p < l.size();) {
         while (p < l.size() && std::isspace((unsigned char)l[p])) ++p;
         if (p >= l.size()) break;
         size_t q = p;
         while (q < l.size() && !std::isspace((unsigned char)l[q])) ++q;
         tok.push_back({p, q - p});
         p = q;
      }
      if ((int64_t)tok.size() == want) {
         for (const auto& t : tok)
            if (!parse(l.c_str() + t.first, (int)t.second, out[k++])) return false;
         continue;
      }
      for (int it = 0; it < fmt.per_line && k < count; ++it) {
         const size_t pos = (size_t)it * fmt.width;
         if (pos >= l.size()) break;
         if (!parse(l.c_str() + pos, (int)std::min<size_t>(fmt.width, l.size() - pos), out[k])) return false;
         ++k;
      }
   }
   return true;
}

// Made-up values of a pattern-only file: off-diagonal entries uniform in
// [-1, 1] from a hash of (seed, row, column), each diagonal entry one more
// than the absolute sum of its row (full symmetric row): symmetric and
// diagonally dominant, hence positive definite.
double made_up(uint64_t seed, int64_t i, int64_t j) {
   uint64_t z = seed ^ ((uint64_t)i * 0x9e3779b97f4a7c15ull) ^ ((uint64_t)j << 32 | (uint64_t)j >> 32);
   z += 0x9e3779b97f4a7c15ull;
   z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
   z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
   z ^= z >> 31;
   return 2.0 * ((double)(z >> 11) * 0x1.0p-53) - 1.0;
}

}  // namespace

extern "C" int sylver_amd_rb_read(const char* filename, int values, unsigned long long seed, int* nrow_out,
                                  int* ncol_out, long** ptr_out, int** row_out, double** val_out,
                                  char* type_out) {
   if (!filename || !nrow_out || !ncol_out || !ptr_out || !row_out || !val_out) return SYLVER_AMD_RB_ERROR_ARGS;
   *ptr_out = nullptr;
   *row_out = nullptr;
   *val_out = nullptr;
   FILE* f = std::fopen(filename, "r");
   if (!f) return SYLVER_AMD_RB_ERROR_OPEN;
   Lines in{f, {}};
   auto fail = [&](int code) {
      std::fclose(f);
      return code;
   };
   if (!in.next()) return fail(SYLVER_AMD_RB_ERROR_FORMAT);  // title + key
   if (!in.next()) return fail(SYLVER_AMD_RB_ERROR_FORMAT);
   long long cards[5] = {0, 0, 0, 0, 0};
   const int ncards = std::sscanf(in.line.c_str(), "%lld %lld %lld %lld %lld", &cards[0], &cards[1], &cards[2],
                                  &cards[3], &cards[4]);
   if (ncards < 3) return fail(SYLVER_AMD_RB_ERROR_FORMAT);
   const long long rhscrd = ncards >= 5 ? cards[4] : 0;
   if (!in.next()) return fail(SYLVER_AMD_RB_ERROR_FORMAT);
   char mt[4] = {0, 0, 0, 0};
   long long nr = 0, nc = 0, nz = 0, ne = 0;
   {
      std::string l = in.line;
      size_t p = 0;
      while (p < l.size() && std::isspace((unsigned char)l[p])) ++p;
      if (p + 3 > l.size()) return fail(SYLVER_AMD_RB_ERROR_FORMAT);
      for (int k = 0; k < 3; ++k) mt[k] = (char)std::tolower((unsigned char)l[p + k]);
      if (std::sscanf(l.c_str() + p + 3, "%lld %lld %lld %lld", &nr, &nc, &nz, &ne) < 3)
         return fail(SYLVER_AMD_RB_ERROR_FORMAT);
   }
   if (type_out) std::memcpy(type_out, mt, 4);
   if (mt[0] == 'c' || mt[2] == 'e') return fail(SYLVER_AMD_RB_ERROR_TYPE);
   if (nr < 0 || nc < 0 || nz < 0 || nr > INT32_MAX || nc > INT32_MAX) return fail(SYLVER_AMD_RB_ERROR_FORMAT);
   if (!in.next()) return fail(SYLVER_AMD_RB_ERROR_FORMAT);
   // formats: parenthesized groups in order
   std::vector<std::string> fmts;
   {
      const std::string& l = in.line;
      size_t p = 0;
      while ((p = l.find('(', p)) != std::string::npos) {
         size_t q = l.find(')', p);
         if (q == std::string::npos) break;
         fmts.push_back(l.substr(p, q - p + 1));
         p = q + 1;
      }
   }
   const bool has_values = mt[0] != 'p' && (ncards < 4 || cards[3] > 0);
   Field fptr, find, fval;
   if (fmts.size() < 2 || !parse_format(fmts[0], fptr) || !parse_format(fmts[1], find))
      return fail(SYLVER_AMD_RB_ERROR_FORMAT);
   if (has_values && (fmts.size() < 3 || !parse_format(fmts[2], fval))) return fail(SYLVER_AMD_RB_ERROR_FORMAT);
   if (rhscrd > 0 && !in.next()) return fail(SYLVER_AMD_RB_ERROR_FORMAT);  // Harwell-Boeing rhs header
   if (!has_values && values == 1) return fail(SYLVER_AMD_RB_ERROR_NO_VALUES);
   std::vector<int64_t> cp, ri;
   std::vector<double> vv;
   if (!read_block(in, fptr, nc + 1, cp, parse_int) || !read_block(in, find, nz, ri, parse_int))
      return fail(SYLVER_AMD_RB_ERROR_FORMAT);
   if (has_values && !read_block(in, fval, nz, vv, parse_real)) return fail(SYLVER_AMD_RB_ERROR_FORMAT);
   std::fclose(f);
   if (cp[0] != 1 || cp[nc] != nz + 1) return SYLVER_AMD_RB_ERROR_FORMAT;
   for (long long j = 0; j < nc; ++j)
      if (cp[j + 1] < cp[j]) return SYLVER_AMD_RB_ERROR_FORMAT;
   for (int64_t r : ri)
      if (r < 1 || r > nr) return SYLVER_AMD_RB_ERROR_FORMAT;
   if (nr != nc) return SYLVER_AMD_RB_ERROR_TYPE;  // the solver needs a square matrix
   const int n = (int)nc;
   const bool symmetric_type = mt[1] == 's' || mt[1] == 'h';
   // lower-triangle entries: (row, col, value), column by column
   std::vector<int64_t> cnt(n + 1, 0);
   for (int j = 0; j < n; ++j)
      for (int64_t p = cp[j] - 1; p < cp[j + 1] - 1; ++p) {
         const int i = (int)ri[p] - 1;
         if (i >= j)
            cnt[j + 1]++;
         else if (symmetric_type)
            cnt[i + 1]++;  // stored above the diagonal: its mirror
      }
   for (int j = 0; j < n; ++j) cnt[j + 1] += cnt[j];
   const int64_t lnz = cnt[n];
   long* optr = static_cast<long*>(std::malloc(sizeof(long) * (n + 1)));
   int* orow = static_cast<int*>(std::malloc(sizeof(int) * std::max<int64_t>(lnz, 1)));
   double* oval = static_cast<double*>(std::malloc(sizeof(double) * std::max<int64_t>(lnz, 1)));
   if (!optr || !orow || !oval) {
      std::free(optr);
      std::free(orow);
      std::free(oval);
      return SYLVER_AMD_RB_ERROR_ALLOC;
   }
   std::vector<int64_t> pos(cnt.begin(), cnt.end() - 1);
   for (int j = 0; j < n; ++j)
      for (int64_t p = cp[j] - 1; p < cp[j + 1] - 1; ++p) {
         const int i = (int)ri[p] - 1;
         int r = i, c = j;
         if (i < j) {
            if (!symmetric_type) continue;
            r = j;
            c = i;
         }
         const int64_t k = pos[c]++;
         orow[k] = r + 1;
         oval[k] = has_values ? vv[p] : 0.0;
      }
   // rows sorted within each column (a mirrored entry may land out of order)
   for (int j = 0; j < n; ++j) {
      const int64_t a = cnt[j], b = cnt[j + 1];
      bool sorted = true;
      for (int64_t k = a + 1; k < b && sorted; ++k) sorted = orow[k - 1] <= orow[k];
      if (sorted) continue;
      std::vector<std::pair<int, double>> e;
      for (int64_t k = a; k < b; ++k) e.push_back({orow[k], oval[k]});
      std::stable_sort(e.begin(), e.end(), [](const auto& x, const auto& y) { return x.first < y.first; });
      for (int64_t k = a; k < b; ++k) {
         orow[k] = e[k - a].first;
         oval[k] = e[k - a].second;
      }
   }
   int flag = 0;
   if (!has_values) {
      // values == 0 leaves zeros (pattern only); otherwise made up (a warning)
      if (values >= 2) {
         std::vector<double> rowsum(n, 0.0);
         for (int j = 0; j < n; ++j)
            for (int64_t k = cnt[j]; k < cnt[j + 1]; ++k) {
               const int i = orow[k] - 1;
               if (i == j) continue;
               const double v = made_up(seed, i, j);
               oval[k] = v;
               rowsum[i] += std::fabs(v);
               rowsum[j] += std::fabs(v);
            }
         for (int j = 0; j < n; ++j)
            for (int64_t k = cnt[j]; k < cnt[j + 1]; ++k)
               if (orow[k] - 1 == j) oval[k] = 1.0 + rowsum[j];
      }
      flag = SYLVER_AMD_RB_WARNING_MADE_UP;
   }
   for (int j = 0; j <= n; ++j) optr[j] = (long)(cnt[j] + 1);
   *nrow_out = n;
   *ncol_out = n;
   *ptr_out = optr;
   *row_out = orow;
   *val_out = oval;
   return flag;
}

extern "C" void sylver_amd_rb_free(void* p) { std::free(p); }
