/* libqdx file formats: QDng's key/value meta files, raw binary data files
 * and the protobuf stream format, so that QDng inputs (potentials, dipoles,
 * lasers, initial wave functions) and QDng's outputs load unchanged, and
 * wave functions and lasers are written the way QDng writes them.
 *
 * Formats (FileSingle, src/tools/FileSingle.h):
 *   - binary: "raw + .meta" (FileSingle.h:1213-1410);
 *   - stream: one file opening with the "QDngPrtoBuf" intro, then per object
 *     a length-prefixed FileSingleHeader message (FileSingleHeader.proto)
 *     and its payload (FileSingle.h:950-1210); wave functions and grid
 *     operators carry a length-prefixed GridSystemHeader (GridSystem.proto)
 *     before their values (WFGridSystem.cpp:140-177, OGridsystem.cpp:63-95),
 *     lasers their size and dt (Laser.cpp:206-235), a multistate wave
 *     function its state count followed by one entry per state
 *     (WFMultistate.cpp:279-293, FileWF.cpp:112-128, 262-276);
 *   - ZLIB compression (FileSingle.h:595-670): a zlib stream of the payload,
 *     in the raw format followed by the tag "ZLIBCOM"; compressed wave
 *     functions hold QDng's lossy k-space reduction (WFGridSystem::Reduce /
 *     Restore, WFGridSystem.cpp:90-137).  BZIP needs libbz2, which is absent:
 *     such files are rejected with QDX_ERR_INCOMPAT.
 * Reading detects the format (FileSingle::DetectFormat, FileSingle.h:
 * 1110-1140) and the compression.  Writing follows qdx_io_set_format
 * (QDng's global parameters wf_format, compress and comptol, FileWF.cpp:
 * 17-56); the library's default is the binary format without compression
 * (QDng writes streams unless wf_format = "binary"; both programs read both).
 *
 * All functions except qdx_wf_load / qdx_wf_save and the qdx_ham_* ones are
 * host-only and need no device.  Errors as in qdx.h; file errors return
 * QDX_ERR_IO (QDng EIOError, src/tools/Exception.h:86-92). */
#ifndef QDX_IO_H
#define QDX_IO_H

#include <stdint.h>

#include "qdx.h"

#ifdef __cplusplus
extern "C" {
#endif

/* ---- output format ---------------------------------------------------------
 * wf_format: QDX_FORMAT_BINARY or QDX_FORMAT_STREAM, used by qdx_wf_save,
 * qdx_laser_save and qdx_gridop_save; compress != 0 writes wave functions
 * ZLIB-compressed after the k-space reduction with tolerance comptol (QDng
 * default 1e-6; comptol <= 0 keeps that).  Process-wide. */
#define QDX_FORMAT_BINARY 0
#define QDX_FORMAT_STREAM 1
int qdx_io_set_format(int32_t wf_format, int32_t compress, double comptol);
int qdx_io_get_format(int32_t *wf_format, int32_t *compress, double *comptol);
/* the format of an existing data file: QDX_FORMAT_STREAM when it opens with
 * the stream intro, else QDX_FORMAT_BINARY (QDX_ERR_IO if it cannot be read) */
int qdx_file_format(const char *path, int32_t *format);

/* ---- key/value meta files (ParamContainer, src/tools/ParamContainer.cpp) ---
 * Reading (Parse, ParamContainer.cpp:731-782): every line holding '=' splits
 * at the first '='; key and value are trimmed; a "[type]" suffix of the key
 * is dropped; lines whose key is empty or starts with '#' are ignored; a
 * repeated key keeps the last value.  Writing (Write, ParamContainer.cpp:
 * 811-832): "key = value" lines in key order (std::map, byte order);
 * doubles as an iostream prints them (6 significant digits,
 * ParamContainer.cpp:584-592). */
typedef struct qdx_meta qdx_meta;
int qdx_meta_create(qdx_meta **out);
/* QDX_ERR_IO if the file cannot be opened (ReadFromFile, ParamContainer.cpp:785-796) */
int qdx_meta_read(const char *path, qdx_meta **out);
int qdx_meta_write(const qdx_meta *m, const char *path);
int qdx_meta_size(const qdx_meta *m, int32_t *n);
/* i-th key in key order; cap: room for the terminating zero (QDX_ERR_PARAM if short) */
int qdx_meta_key(const qdx_meta *m, int32_t i, char *key, int32_t cap);
/* QDX_ERR_PARAM if the key is absent */
int qdx_meta_get(const qdx_meta *m, const char *key, char *value, int32_t cap);
int qdx_meta_set(qdx_meta *m, const char *key, const char *value);
int qdx_meta_set_int(qdx_meta *m, const char *key, int64_t value);
int qdx_meta_set_double(qdx_meta *m, const char *key, double value);
int qdx_meta_destroy(qdx_meta *m);

/* ---- raw data files (FileSingle<C> binary format) -------------------------
 * name: path without suffix.  counter < 0: the data file is name + suffix;
 * counter >= 0 (a sequence, FileSingle.h:1398-1410): name_<counter> + suffix.
 * The meta file is always name.meta. */

/* Meta data of a file (ReadMetaFromRaw, FileSingle.h:1217-1238): name.meta,
 * else, when name holds '_', the name cut at its last '_' + ".meta". */
int qdx_file_meta(const char *name, qdx_meta **out);
/* Read exactly `bytes` bytes of the data file (ReadDataFromRaw,
 * FileSingle.h:1245-1287: a shorter file is an error, a longer one is read
 * from its start); with a ZLIB tag in the file's last 8 bytes
 * (DetectCompressionInPayload, FileSingle.h:769-801) the inflated data
 * (BZIP: QDX_ERR_INCOMPAT). */
int qdx_file_read(const char *name, const char *suffix, int32_t counter, void *dst, int64_t bytes);
/* Write name.meta (when meta is not NULL) and the data file
 * (WriteSingleFileToRaw, FileSingle.h:1303-1360, uncompressed). */
int qdx_file_write(const char *name, const char *suffix, int32_t counter, const qdx_meta *meta, const void *src,
                   int64_t bytes);

/* ---- grid operators, lasers, propagation meta ----------------------------- */

/* Grid data of an operator file name.op + meta (OGridPotential::Init with
 * "file", OGridPotential.cpp:36-80; the grid keys dims, N<d>, xmin<d>,
 * xmax<d> of GridSystemHeaderPm, GridSystem.cpp:17-53).  setzero != 0
 * subtracts the minimum, then offset is added, then the values are
 * multiplied by scale (OGridPotential.cpp:55-76).  data: room for cap
 * doubles (prod N); *n receives prod N.  data may be NULL to query the grid. */
int qdx_gridop_load(const char *name, int32_t setzero, double offset, double scale, qdx_grid *grid, double *data,
                    int64_t cap, int64_t *n);
/* Write an operator file (class GridPotential) in the same layout. */
int qdx_gridop_save(const char *name, const qdx_grid *grid, const double *data);

/* Laser field name.op + meta (Laser::Init, Laser.cpp:35-50: keys dt, Nt;
 * OLaser::Init, OLaser.cpp:133-160: suffix .op).  data may be NULL to query
 * dt and nt. */
int qdx_laser_load(const char *name, double *dt, int64_t *nt, double *data, int64_t cap);
int qdx_laser_save(const char *name, double dt, int64_t nt, const double *data);

/* Propagation.meta in dir (ProgPropa::_WriteMeta, ProgPropa.cpp:144-158):
 * CLASS = Propagation, Nt, dt, WFBaseName, Wcycle.  dir: "" or ending in '/'. */
int qdx_propagation_meta(const char *dir, int64_t nt, double dt, const char *wf_base, int32_t wcycle);

/* ---- device objects from files -------------------------------------------- */

/* Wave function from files (FileWF::LoadWaveFunctionByMeta, FileWF.cpp:67-131):
 * binary format: class WFGridCartesian (one state) or WFMultistate (key
 * "states"; state i in name-<i>, its own meta file name-<i>.meta, all on one
 * grid); stream format: name[_counter].wf holding one WFGridCartesian or a
 * WFMultistate with its states; compressed data restored from the k-space
 * reduction.  counter as for qdx_file_read. */
int qdx_wf_load(qdx_ctx *ctx, const char *name, int32_t counter, qdx_wf **out);
/* Write a wave function (FileWF::WriteWaveFunction, FileWF.cpp:231-283) in
 * the format of qdx_io_set_format.  Binary: one state as name[_counter].wf +
 * name.meta (class WFGridCartesian); several as name-<i>[_counter].wf +
 * name-<i>.meta per state and name.meta (class WFMultistate, states).
 * Stream: name[_counter].wf holding everything. */
int qdx_wf_save(qdx_wf *wf, const char *name, int32_t counter);

/* Hamiltonian terms read from QDng files: a potential (qdx_ham_add_potential)
 * from name.op with setzero / offset / scale, and a dipole
 * (qdx_ham_add_dipole) from mu_name.op and the laser file laser_name.op; as in
 * QDng, a GridDipole's scale multiplies both the dipole array
 * (OGridPotential::Init) and the field (OLaser::Clock, Laser::Get).
 * The file's grid must equal the Hamiltonian's (QDX_ERR_INCOMPAT). */
int qdx_ham_add_potential_file(qdx_ham *h, int32_t row, int32_t col, const char *name, int32_t setzero,
                               double offset, double scale);
int qdx_ham_add_dipole_file(qdx_ham *h, int32_t row, int32_t col, const char *mu_name, const char *laser_name,
                            double scale);

#ifdef __cplusplus
}
#endif

#endif /* QDX_IO_H */
