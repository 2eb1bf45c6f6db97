/*
 * Image files of the bm4d-of, tvl1flow and tvl1flow-seq command-line tools
 * (libbm4dof_io.so, host code only).  They replace the image calls the
 * reference's tools make through iio:
 *   iio_read_image_float_split   lib/iio/iio.c:3242  (bm4dof_io_read)
 *   iio_save_image_float_split   lib/iio/iio.c:3870  (bm4dof_io_write)
 * for the formats its hot path uses: PNG (8/16-bit read, 8-bit write with
 * the reference's quantisation (uint8)clamp(v + 0.5, 0, 255)), uncompressed
 * TIFF (8/16-bit unsigned or 32-bit float read, 32-bit float write) and
 * Middlebury .flo (2 channels).  Images are planar float:
 * data[c*w*h + y*w + x].
 */
#ifndef BM4DOF_IO_H_INCLUDED
#define BM4DOF_IO_H_INCLUDED

#ifdef __cplusplus
extern "C" {
#endif

/* Reads a PNG, TIFF or .flo file (format from its first bytes).  *data is
 * allocated with malloc (free with bm4dof_io_free).  Returns 0, or -1 with
 * bm4dof_io_error() describing the failure. */
int bm4dof_io_read(const char *path, float **data, int *w, int *h, int *c);
/* One gray channel as tvl1flow reads its frames (iio_read_image_float,
 * lib/iio/iio.c:3486, via lib/tvl1flow/main.c:42-49): 3 or 4 channels give
 * .299 R + .587 G + .114 B in double, stored in the file's sample type
 * (integer samples truncate); a 2-channel file is an error. */
int bm4dof_io_read_gray(const char *path, float **data, int *w, int *h);
/* Writes by extension: .png, .tif/.tiff or .flo.  Returns 0 or -1. */
int bm4dof_io_write(const char *path, const float *data, int w, int h, int c);
void bm4dof_io_free(float *data);
const char *bm4dof_io_error(void);

#ifdef __cplusplus
}
#endif

#endif /* BM4DOF_IO_H_INCLUDED */
