/*
 * orc_main.c -- command-line runner of the oracle's whole dana run.
 * TEST INFRASTRUCTURE ONLY (golden-frame checks and the CPU baseline).
 *
 * usage: orc_run <dir>   reads <dir>/entrada.ini, <dir>/movedor.ini and, for
 *        the chunks reservoir, <dir>/chunk.xyz (the formats of
 *        dana.F90:309-327, 399-427, 552-587); prints the final frame the way
 *        salida() does (sym x y z z) preceded by "nat zmax".
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "orc.h"

static int read_num(FILE *f, double *v) {
  char line[512];
  if (!fgets(line, sizeof line, f)) return -1;
  return sscanf(line, "%lf", v) == 1 ? 0 : -1;
}

int main(int argc, char **argv) {
  if (argc < 2) {
    fprintf(stderr, "usage: %s <dir> [nst]\n", argv[0]);
    return 2;
  }
  char path[1024];
  snprintf(path, sizeof path, "%s/entrada.ini", argv[1]);
  FILE *f = fopen(path, "r");
  if (!f) { perror(path); return 2; }
  double v[13];
  for (int i = 0; i < 13; i++)
    if (read_num(f, &v[i])) { fprintf(stderr, "bad entrada line %d\n", i + 1); return 2; }
  fclose(f);
  orc_run_cfg c;
  memset(&c, 0, sizeof c);
  c.idum = (int)v[0]; c.prob = v[1]; c.h = v[2]; c.nst = (int)v[3]; c.nwr = (int)v[4];
  c.xi = v[5]; c.yi = v[6]; c.dist = v[7]; c.z0 = v[8]; c.zmax = v[9];
  c.dif_sc = v[10]; c.dif_sei = v[11]; c.nb_dcut = v[12];
  if (argc > 2) c.nst = atoi(argv[2]);
  snprintf(path, sizeof path, "%s/movedor.ini", argv[1]);
  f = fopen(path, "r");
  if (!f) { perror(path); return 2; }
  char line[512], word[128];
  if (!fgets(line, sizeof line, f)) return 2;
  c.ermak = strstr(line, ".true.") || strstr(line, ".TRUE.") || strstr(line, "T") == line;
  if (!fgets(line, sizeof line, f)) return 2;
  sscanf(line, "%127s", word);
  if (!strcmp(word, "piston")) c.reservoir = 0;
  else if (!strcmp(word, "chunks")) c.reservoir = 1;
  else if (!strcmp(word, "gcmc")) c.reservoir = 2;
  else c.reservoir = -1;
  if (c.reservoir == 2) {
    if (!fgets(line, sizeof line, f)) return 2;
    sscanf(line, "%lf %d", &c.act, &c.nadj);
  }
  fclose(f);
  double *chunk = NULL;
  if (c.reservoir == 1) {
    snprintf(path, sizeof path, "%s/chunk.xyz", argv[1]);
    f = fopen(path, "r");
    if (!f) { perror(path); return 2; }
    if (!fgets(line, sizeof line, f)) return 2;
    c.nchunk = atoi(line);
    if (!fgets(line, sizeof line, f)) return 2;
    chunk = (double *)malloc(sizeof(double) * 3 * (size_t)c.nchunk);
    for (int j = 0; j < c.nchunk; j++) {
      char sym[16];
      if (!fgets(line, sizeof line, f)) return 2;
      sscanf(line, "%15s %lf %lf %lf", sym, &chunk[3 * j], &chunk[3 * j + 1], &chunk[3 * j + 2]);
    }
    fclose(f);
    c.chunk_pos = chunk;
  }
  orc_run_out o;
  memset(&o, 0, sizeof o);
  o.cap = 1 << 22;
  o.pos = (double *)malloc(sizeof(double) * 3 * (size_t)o.cap);
  o.z = (int8_t *)malloc((size_t)o.cap);
  int st = orc_dana_run(&c, &o);
  if (st) { fprintf(stderr, "orc_dana_run failed: %d\n", st); return 1; }
  printf("%d %.17g\n", o.nat, o.zmax);
  for (int i = 0; i < o.nat; i++)
    printf("%s %.17g %.17g %.17g %d\n", o.z[i] == 2 ? "CG" : (o.z[i] == 3 ? "F" : "Li"),
           o.pos[3 * i], o.pos[3 * i + 1], o.pos[3 * i + 2], o.z[i]);
  fprintf(stderr, "loop %.3f s, particle-steps %lld, rate %.3e /s, nupd %lld choques %lld try %lld depo %lld ins %lld del %lld\n",
          o.loop_seconds, (long long)o.particle_steps,
          o.particle_steps / (o.loop_seconds > 0 ? o.loop_seconds : 1e-9),
          (long long)o.c.nupd, (long long)o.c.choques, (long long)o.c.try_,
          (long long)o.c.depo, (long long)o.c.ins, (long long)o.c.del);
  free(chunk);
  return 0;
}
