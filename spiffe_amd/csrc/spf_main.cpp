// spf_main.cpp -- command-line runner: spiffe-amd <deck> [-max_particles N]
// [-max_steps N] [-quiet] [-fixed].  Runs the deck on GPU 0 with
// libspiffe_amd (spf_deck_run) and prints a one-line summary.
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "../../include/spiffe_amd.h"

int main(int argc, char **argv) {
  if (argc < 2) {
    fprintf(stderr, "usage: %s <deck> [-max_particles N] [-max_steps N] [-quiet] [-fixed]\n", argv[0]);
    return 1;
  }
  spf_deck_options o;
  spf_deck_defaults(&o);
  o.verbose = 1;
  const char *deck = nullptr;
  for (int i = 1; i < argc; i++) {
    if (!strcmp(argv[i], "-max_particles") && i + 1 < argc) o.max_particles = atoll(argv[++i]);
    else if (!strcmp(argv[i], "-max_steps") && i + 1 < argc) o.max_steps = atoll(argv[++i]);
    else if (!strcmp(argv[i], "-quiet")) o.verbose = 0;
    else if (!strcmp(argv[i], "-fixed")) o.compat = 0;
    else deck = argv[i];
  }
  spf_deck_stats st;
  char err[1024] = {0};
  int rc = spf_deck_run(deck, &o, &st, err, sizeof(err));
  if (rc) {
    fprintf(stderr, "spiffe-amd: %s (code %d)\n", err, rc);
    return rc;
  }
  printf("steps %lld, particles %lld, emitted %lld, screen rows %lld, lost %lld, field saves %lld\n",
         (long long)st.steps, (long long)st.particles, (long long)st.emitted, (long long)st.screen_rows,
         (long long)st.lost_rows, (long long)st.field_saves);
  return 0;
}
