#pragma once
#include <string>

#define TABLE_MAXN 6

namespace pypde {

struct Tables {
  int N;
  double x[TABLE_MAXN], w[TABLE_MAXN], end0[TABLE_MAXN], end1[TABLE_MAXN];
  double der[TABLE_MAXN][TABLE_MAXN];
  double wmat[4][TABLE_MAXN][TABLE_MAXN], winv[4][TABLE_MAXN][TABLE_MAXN];
  double sig[TABLE_MAXN][TABLE_MAXN];
  double dgmat[TABLE_MAXN][TABLE_MAXN], dginv[TABLE_MAXN][TABLE_MAXN];
  double flc[TABLE_MAXN], flm[TABLE_MAXN][TABLE_MAXN][TABLE_MAXN];
  int nsten, soff[4];
  double slam[4];
};

int make_tables(int n, Tables* T);
std::string tables_source(const Tables& T, int D);

}  // namespace pypde
