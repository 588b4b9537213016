/*
 * plk_align.h -- alignment ingest for libplk: PHYLIP / FASTA reading and site-pattern
 * compression, producing the states / ptn_freq arrays plk_set_patterns takes.
 *
 * Follows IQ-TREE's Alignment:
 *   PHYLIP            doReadPhylip alignment/alignment.cpp:2474-2556 (interleaved; names cut at
 *                     the first blank, else 10 characters), doReadPhylipSequential :2567-2633
 *   FASTA             doReadFasta :2708-2808 (names shortened word by word while unique)
 *   sequence text     processSeq :2444-2472 (blanks skipped, upper-cased, '!' a gap,
 *                     (..)/{..} one unknown character)
 *   type detection    detectSequenceType :1720-1786 (DNA / protein)
 *   state codes       buildStateMap :1788-1846 (DNA ambiguity codes, protein B/Z/J,
 *                     gaps and unknowns -> STATE_UNKNOWN)
 *   codons            buildPattern :2219-2330 (standard code; stop codons and partly
 *                     ambiguous codons -> STATE_UNKNOWN)
 *   patterns          addPatternLazy :1511-1543 (first-occurrence order, frequencies,
 *                     site -> pattern map), computeConst :1321-1420 (constant flag)
 * Errors (unreadable file, malformed header, unequal lengths, invalid characters,
 * duplicated names) return PLK_ERR_ARG with the reference's message in plk_last_error().
 */
#ifndef PLK_ALIGN_H
#define PLK_ALIGN_H

#include <stddef.h>
#include <stdint.h>

#include "plk.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct plk_align plk_align;

/* PLK_ALIGN_DETECT, or PLK_SEQ_DNA / PLK_SEQ_PROTEIN / PLK_SEQ_CODON (plk_model.h). */
#define PLK_ALIGN_DETECT (-1)
/* format: 0 = detect ('>' first -> FASTA, else PHYLIP), 1 = PHYLIP (interleaved),
 * 2 = PHYLIP sequential, 3 = FASTA */
int plk_align_read(plk_align **out, const char *path, int format, int seq_type);
int plk_align_parse(plk_align **out, const char *text, size_t len, int format, int seq_type);
int plk_align_destroy(plk_align *a);

/* dims[0..5]: ntaxa, nsites (alignment columns; codons for the codon type), npatterns,
 * nstates (4, 20, 61), seq_type, sites made only of gaps/unknowns. */
int plk_align_dims(const plk_align *a, int32_t *dims);
/* Taxon name (NUL-terminated, truncated to buflen). */
int plk_align_name(const plk_align *a, int taxon, char *buf, int buflen);
/* states: ntaxa x npatterns (row per taxon, STATE_UNKNOWN = plk_state_unknown()),
 * freq: npatterns, site_pattern: nsites, flags: npatterns (bit 0 constant pattern,
 * bit 1 parsimony-informative).  Any pointer may be NULL. */
int plk_align_patterns(const plk_align *a, uint8_t *states, double *freq, int32_t *site_pattern, int32_t *flags);

#ifdef __cplusplus
}
#endif
#endif
