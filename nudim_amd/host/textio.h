/*
 * textio -- the text files of NUDIM's command-line tools, read and written
 * with several host threads.
 *
 * The reference reads one value per fscanf("%le\n") (src/reconstruction.c:
 * 341-358, src/projection_sim.c:257-266) and writes one per fprintf
 * ("%le\n" / "%f\n", src/reconstruction.c:531-540, 707-719,
 * src/projection_sim.c:396-403).  Here a file is read whole, cut at
 * whitespace into one segment per thread and parsed with strtod (the
 * conversion fscanf's %le performs, correctly rounded), and output ranges
 * are formatted with snprintf in parallel and written in order, so files
 * are byte for byte those of the sequential calls.
 */
#ifndef NUDIM_TEXTIO_H
#define NUDIM_TEXTIO_H

#ifdef __cplusplus
extern "C" {
#endif

/* Parse up to `count` whitespace-separated doubles of the file into out.
 * Returns the number parsed (< count for a short file, stopping at the
 * first token strtod rejects), or -1 when the file cannot be opened. */
long textio_read_doubles(const char *path, double *out, long count, int nthreads);

/* Write count values, each formatted with fmt (one conversion, e.g. "%f\n").
 * Returns 0, or -1 when the file cannot be written. */
int textio_write_doubles(const char *path, const double *v, long count, const char *fmt, int nthreads);

/* Threads to use: NUDIM_IO_THREADS if set, else the online CPUs, at most 16. */
int textio_threads(void);

#ifdef __cplusplus
}
#endif

#endif
