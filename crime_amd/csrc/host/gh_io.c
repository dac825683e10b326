/*
 * Output files of the GetHI host side: HEALPix FITS maps and the frequency table.
 *
 * write_maps / write_single_fit / write_nu_table (src/io_gh.c:69-131) and he_write_healpix_map
 * (src/healpix_extra.c:132-164).  The FITS file is written directly (no cfitsio): an empty primary
 * HDU and a BINTABLE extension with one 1E column "T" in mK and the keywords PIXTYPE=HEALPIX,
 * ORDERING=RING, NSIDE and COORDSYS=G, data big-endian, every unit padded to 2880 bytes
 * (FITS Standard 4.0).
 */
#include <stdarg.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "gethi_host.h"

#define FITS_BLOCK 2880

typedef struct { char buf[FITS_BLOCK * 4]; int n; } fits_hdr;

static void card(fits_hdr *h, const char *fmt, ...)
{
  char c[81];
  va_list ap;
  va_start(ap, fmt);
  int len = vsnprintf(c, sizeof(c), fmt, ap);
  va_end(ap);
  if (len > 80) len = 80;
  memset(h->buf + h->n, ' ', 80);
  memcpy(h->buf + h->n, c, (size_t)len);
  h->n += 80;
}

static void card_str(fits_hdr *h, const char *key, const char *val, const char *comment)
{
  char q[72];
  /* string values are quoted and padded to at least 8 characters */
  snprintf(q, sizeof(q), "'%-8s'", val);
  if (comment) card(h, "%-8s= %-20s / %s", key, q, comment);
  else card(h, "%-8s= %-20s", key, q);
}

static void card_int(fits_hdr *h, const char *key, long v, const char *comment)
{
  if (comment) card(h, "%-8s= %20ld / %s", key, v, comment);
  else card(h, "%-8s= %20ld", key, v);
}

static void card_bool(fits_hdr *h, const char *key, int v, const char *comment)
{
  card(h, "%-8s= %20s / %s", key, v ? "T" : "F", comment);
}

static int hdr_flush(fits_hdr *h, FILE *f)
{
  card(h, "END");
  int pad = (FITS_BLOCK - h->n % FITS_BLOCK) % FITS_BLOCK;
  memset(h->buf + h->n, ' ', (size_t)pad);
  h->n += pad;
  return fwrite(h->buf, 1, (size_t)h->n, f) == (size_t)h->n ? 0 : -1;
}

int gh_write_healpix_fits(const char *fname, const float *map, long n_side)
{
  long npix = 12 * n_side * n_side;
  FILE *f = fopen(fname, "wb");
  if (!f) return -1;
  int rc = 0;
  fits_hdr h;
  h.n = 0;
  card_bool(&h, "SIMPLE", 1, "file does conform to FITS standard");
  card_int(&h, "BITPIX", 8, "number of bits per data pixel");
  card_int(&h, "NAXIS", 0, "number of data axes");
  card_bool(&h, "EXTEND", 1, "FITS dataset may contain extensions");
  rc |= hdr_flush(&h, f);

  h.n = 0;
  card_str(&h, "XTENSION", "BINTABLE", "binary table extension");
  card_int(&h, "BITPIX", 8, "8-bit bytes");
  card_int(&h, "NAXIS", 2, "2-dimensional binary table");
  card_int(&h, "NAXIS1", 4, "width of table in bytes");
  card_int(&h, "NAXIS2", npix, "number of rows in table");
  card_int(&h, "PCOUNT", 0, "size of special data area");
  card_int(&h, "GCOUNT", 1, "one data group (required keyword)");
  card_int(&h, "TFIELDS", 1, "number of fields in each row");
  card_str(&h, "TTYPE1", "T", "label for field   1");
  card_str(&h, "TFORM1", "1E", "data format of field: 4-byte REAL");
  card_str(&h, "TUNIT1", "mK", "physical unit of field");
  card_str(&h, "EXTNAME", "BINTABLE", "name of this binary table extension");
  card_str(&h, "PIXTYPE", "HEALPIX", "HEALPIX Pixelisation");
  card_str(&h, "ORDERING", "RING", "Pixel ordering scheme, either RING or NESTED");
  card_int(&h, "NSIDE", n_side, "Resolution parameter for HEALPIX");
  card_str(&h, "COORDSYS", "G", "Pixelisation coordinate system");
  card(&h, "COMMENT G = Galactic, E = ecliptic, C = celestial = equatorial");
  rc |= hdr_flush(&h, f);

  /* big-endian data in chunks */
  enum { CH = 1 << 16 };
  uint32_t *tmp = (uint32_t *)malloc(CH * sizeof(uint32_t));
  if (!tmp) {
    fclose(f);
    return -1;
  }
  for (long i0 = 0; i0 < npix && !rc; i0 += CH) {
    long n = npix - i0 < CH ? npix - i0 : CH;
    for (long i = 0; i < n; i++) {
      uint32_t u;
      memcpy(&u, map + i0 + i, 4);
      tmp[i] = __builtin_bswap32(u);
    }
    if (fwrite(tmp, 4, (size_t)n, f) != (size_t)n) rc = -1;
  }
  free(tmp);
  long nbytes = npix * 4;
  long pad = (FITS_BLOCK - nbytes % FITS_BLOCK) % FITS_BLOCK;
  if (pad && !rc) {
    char z[FITS_BLOCK];
    memset(z, 0, sizeof(z));
    if (fwrite(z, 1, (size_t)pad, f) != (size_t)pad) rc = -1;
  }
  if (fclose(f)) rc = -1;
  return rc;
}

int gh_write_nutable(const gh_params *p, const char *prefix)
{
  char fname[600];
  snprintf(fname, sizeof(fname), "%s_nuTable.dat", prefix);
  FILE *f = fopen(fname, "w");
  if (!f) return -1;
  for (int ii = 0; ii < p->n_nu; ii++) {
    double nu0 = p->nu0[ii], nuf = p->nuf[ii];
    double z0 = GH_NU_21 / nuf - 1;
    double zf = GH_NU_21 / nu0 - 1;
    fprintf(f, "%d %lf %lf %lf %lf\n", ii + 1, nu0, nuf, z0, zf);
  }
  return fclose(f) ? -1 : 0;
}

int gh_write_maps(const gh_params *p, const float *maps)
{
  if (gh_write_nutable(p, p->prefix_out)) return -1;
  long npix = 12 * p->n_side * p->n_side;
  for (int inu = 0; inu < p->n_nu; inu++) {
    char fname[600];
    snprintf(fname, sizeof(fname), "%s_%03d.fits", p->prefix_out, inu + 1);
    if (gh_write_healpix_fits(fname, maps + (size_t)npix * inu, p->n_side)) return -1;
  }
  return 0;
}
