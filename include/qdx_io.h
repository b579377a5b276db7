/* libqdx file formats: QDng's key/value meta files and raw binary data files,
 * so that QDng inputs (potentials, dipoles, lasers, initial wave functions)
 * load unchanged and wave functions are written the way QDng writes them.
 *
 * Scope: the binary ("raw + .meta") format of FileSingle
 * (src/tools/FileSingle.h:1213-1410), which QDng reads and writes by default
 * (FileWF.h:11: no compression; FileWF.cpp:44-56: wf_format "binary").  Not
 * built: the protobuf stream format (FileSingle.h:950-1210, "QDngPrtoBuf"
 * intro) and compressed payloads (ZLIB/BZIP, FileSingle.h:592-800); a raw
 * file whose last 8 bytes carry a compression tag is rejected with
 * QDX_ERR_INCOMPAT.
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
 * from its start); a compression tag in the file's last 8 bytes
 * (DetectCompressionInPayload, FileSingle.h:769-801) gives QDX_ERR_INCOMPAT. */
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

/* Wave function from files (FileWF::LoadWaveFunctionByMeta, FileWF.cpp:67-131,
 * binary format): class WFGridCartesian (one state) or WFMultistate (key
 * "states"; state i in name-<i>, its own meta file name-<i>.meta, all on one
 * grid).  counter as for qdx_file_read. */
int qdx_wf_load(qdx_ctx *ctx, const char *name, int32_t counter, qdx_wf **out);
/* Write a wave function (FileWF::WriteWaveFunction, FileWF.cpp:231-283,
 * binary format): one state as name[_counter].wf + name.meta (class
 * WFGridCartesian); several as name-<i>[_counter].wf + name-<i>.meta per state
 * and name.meta (class WFMultistate, states). */
int qdx_wf_save(qdx_wf *wf, const char *name, int32_t counter);

/* Hamiltonian terms read from QDng files: a potential (qdx_ham_add_potential)
 * from name.op with setzero / offset / scale, and a dipole
 * (qdx_ham_add_dipole) from mu_name.op and the laser file laser_name.op.
 * The file's grid must equal the Hamiltonian's (QDX_ERR_INCOMPAT). */
int qdx_ham_add_potential_file(qdx_ham *h, int32_t row, int32_t col, const char *name, int32_t setzero,
                               double offset, double scale);
int qdx_ham_add_dipole_file(qdx_ham *h, int32_t row, int32_t col, const char *mu_name, const char *laser_name,
                            double scale);

#ifdef __cplusplus
}
#endif

#endif /* QDX_IO_H */
