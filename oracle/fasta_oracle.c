/*
 * fasta_oracle.c -- CPU oracle of the FASTA normalisation.
 *
 * TEST INFRASTRUCTURE ONLY (see pastis_oracle.h).  Restates what one process
 * of the reference does to a FASTA file before any sequence reaches the
 * aligner:
 *   - DistFastaData::read_fasta (src/DistFastaData.cpp:161-263) with one
 *     rank reads bytes [0, size-2]: the file's last byte is dropped;
 *   - the FastaData constructor (src/FastaData.cpp:13-148) compacts that
 *     buffer in place: newlines inside a sequence are removed except the
 *     one before the next '>', '*' is removed, U/u become T, a '>' starts a
 *     header only outside a header, and a sequence shorter than k is removed
 *     together with its header -- the removal takes out one byte more than
 *     the record (the byte before its '>'), so the sequence before a removed
 *     record that is followed by another record loses its last residue;
 *   - get_sequence (inc/FastaData.hpp:105-129): sequence i spans
 *     [seq_start_i, id_start_{i+1} - 1), the last one [seq_start, l_end + 1),
 *     and its length is taken as a ushort (mod 65536).
 * The restatement follows the constructor's loop over a copy of the buffer
 * with room in front: when the first record is removed the reference writes
 * one byte before its buffer (undefined behaviour), one more for every
 * further record removed at the start; here those bytes land in the room and
 * the other records are unaffected.
 */
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "pastis_oracle.h"

int or_fasta_normalize(const char *file, uint64_t size, int k, char **res_out,
                       uint64_t **offs_out, uint64_t **orig_out,
                       uint64_t *nseq_out)
{
    *res_out = NULL;
    *offs_out = NULL;
    *orig_out = NULL;
    *nseq_out = 0;
    if (size < 2)
        return 0;
    const uint64_t n = size - 1;   /* bytes [0, size-2] */
    /* each erased record at the start reaches one byte further before the
     * buffer; there are fewer records than bytes */
    char *room = (char *)malloc(2 * n + 2);
    char *buff = room + n + 1;
    memcpy(buff, file, n);
    buff[n] = 0;
    /* signed positions: a removed first record moves one byte before 0 */
    int64_t *id_starts = (int64_t *)malloc(sizeof(int64_t) * (n + 1));
    int64_t *seq_starts = (int64_t *)malloc(sizeof(int64_t) * (n + 1));
    uint64_t *orig = (uint64_t *)malloc(sizeof(uint64_t) * (n + 1));
    uint64_t nid = 0, nss = 0, nrec = 0;
    int in_name = 0, in_seq = 0;
    int64_t seq_len = 0, nc = 0;
    const int64_t l_end = (int64_t)n - 1;
    for (int64_t i = 0; i <= l_end; ++i) {
        const char c = buff[i];
        const int64_t idx = i - nc;
        buff[idx] = c;
        if ((c == 'U' || c == 'u') && in_seq)
            buff[idx] = 'T';
        if (c == '>' && !in_name) {
            id_starts[nid] = idx;
            orig[nid] = nrec++;
            ++nid;
            seq_len = 0;
            in_name = 1;
            in_seq = 0;
        } else if (c == '\n') {
            if (in_name && i + 1 <= l_end) {
                seq_starts[nss++] = idx + 1;
                in_name = 0;
                in_seq = 1;
            } else if (in_seq && i + 1 <= l_end) {
                if (buff[i + 1] != '>') {
                    ++nc;
                } else if (seq_len < k) {
                    ++seq_len;
                    nc += seq_len + (seq_starts[nss - 1] - id_starts[nid - 1]) + 1;
                    --nid;
                    --nss;
                }
            }
        } else if (c == '*') {
            if (in_seq)
                ++nc;
        } else if (in_seq) {
            ++seq_len;
        }
    }
    /* a file that ends inside a header line leaves a record without a
     * sequence; the reference then pops mismatched entries (undefined), so
     * the removal is only restated for a complete last record */
    if (seq_len < k && nid > 0 && nss == nid) {
        nc += seq_len + (seq_starts[nss - 1] - id_starts[nid - 1]) + 1;
        --nid;
        --nss;
    }
    const int64_t l_end2 = l_end - nc;
    const uint64_t ns = nss < nid ? nss : nid;
    uint64_t *offs = (uint64_t *)malloc(sizeof(uint64_t) * (ns + 1));
    uint64_t tot = 0;
    offs[0] = 0;
    for (uint64_t s = 0; s < ns; ++s) {
        const int64_t end_pos = s + 1 < nid ? id_starts[s + 1] - 1 : l_end2 + 1;
        /* erased records eating past this sequence's start: no residues
         * (the reference reads a wrapped length there) */
        const int64_t span = end_pos > seq_starts[s] ? end_pos - seq_starts[s] : 0;
        const uint16_t len = (uint16_t)span;
        tot += len;
        offs[s + 1] = tot;
    }
    char *res = (char *)malloc(tot ? tot : 1);
    for (uint64_t s = 0; s < ns; ++s)
        memcpy(res + offs[s], buff + seq_starts[s], offs[s + 1] - offs[s]);
    free(room);
    free(id_starts);
    free(seq_starts);
    *res_out = res;
    *offs_out = offs;
    *orig_out = orig;
    *nseq_out = ns;
    return 0;
}
