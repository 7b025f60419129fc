/* Molecular file front end of the MI355X voronota-ltr library: PDB / mmCIF /
 * XYZR parsing, radii assignment, VMD-like selections and groupings, the
 * from-file tessellation, the CLI's JSON document and the PyMOL CGO script.
 * Host code; only vltr_compute_tessellation_from_file and
 * vltr_pymol_graphics run the device pipeline (through the context).
 * Errors: VLTR_ERR_INVALID_ARGUMENT with the reference's message text
 * (vltr_last_error), VLTR_ERR_IO for files that cannot be read. */
#ifndef VORONOTA_LTR_INPUT_H
#define VORONOTA_LTR_INPUT_H

#include "voronota_ltr.h"

#ifdef __cplusplus
extern "C" {
#endif

#define VLTR_ERR_IO 5

typedef struct vltr_radii vltr_radii;
typedef struct vltr_structure vltr_structure;

/* ParseOptions (src/input/mod.rs:69-77); all zero = the reference's default. */
typedef struct vltr_parse_options {
  int exclude_heteroatoms;
  int include_hydrogens;
  int as_assembly;
} vltr_parse_options;

/* InputFormat (src/input/mod.rs:58-66); VLTR_FORMAT_AUTO detects from content. */
#define VLTR_FORMAT_AUTO (-1)
#define VLTR_FORMAT_PDB 0
#define VLTR_FORMAT_MMCIF 1
#define VLTR_FORMAT_XYZR 2

/* RadiiLookup (src/input/radii.rs:15-111): new() with the default rules,
 * empty(), load_from_text(), get_radius(). */
int vltr_radii_create(int with_default_rules, vltr_radii** out);
int vltr_radii_load_text(vltr_radii* r, const char* text);
double vltr_radii_get(const vltr_radii* r, const char* res_name, const char* atom_name);
void vltr_radii_free(vltr_radii* r);

/* parse_file_with_records (src/input/mod.rs:300-334): format from the file
 * extension, else from the first line. radii may be NULL (default rules),
 * options may be NULL (defaults). */
int vltr_parse_file(const char* path, const vltr_parse_options* options, const vltr_radii* radii,
                    vltr_structure** out);
/* parse_reader_with_records on an in-memory text (stdin in the CLI); format
 * VLTR_FORMAT_AUTO detects it from the first line as the CLI's stdin path
 * does (src/main.rs:206-224: anything not mmCIF or PDB is read as XYZR). */
int vltr_parse_text(const char* text, size_t len, int format, const vltr_parse_options* options,
                    const vltr_radii* radii, vltr_structure** out);
int vltr_structure_format(const vltr_structure* s);
size_t vltr_structure_num_balls(const vltr_structure* s);
const vltr_ball* vltr_structure_balls(const vltr_structure* s);
/* AtomRecord (src/input/mod.rs:27-55): none for XYZR input. */
size_t vltr_structure_num_records(const vltr_structure* s);
#define VLTR_REC_RECORD_NAME 0
#define VLTR_REC_NAME 1
#define VLTR_REC_ALT_LOC 2
#define VLTR_REC_RES_NAME 3
#define VLTR_REC_CHAIN_ID 4
#define VLTR_REC_I_CODE 5
#define VLTR_REC_ELEMENT 6
const char* vltr_record_string(const vltr_structure* s, size_t i, int field);
/* serial, res_seq; xyzb = x, y, z, b_factor */
int vltr_record_numbers(const vltr_structure* s, size_t i, int32_t* serial, int32_t* res_seq, double* xyzb);
void vltr_structure_free(vltr_structure* s);

/* build_chain_grouping / build_residue_grouping (src/input/mod.rs:377-420):
 * one group id per record into out (num_records entries). */
int vltr_grouping_chain(const vltr_structure* s, int32_t* out);
int vltr_grouping_residue(const vltr_structure* s, int32_t* out);
/* parse_selections + build_custom_grouping (src/input/selection.rs:742-814). */
int vltr_grouping_selections(const vltr_structure* s, const char* const* selections, size_t n_selections,
                             int32_t* out);
/* Selection::parse + matches (src/input/selection.rs:721-738): one byte per
 * record, 1 where the selection matches. */
int vltr_selection_match(const vltr_structure* s, const char* selection, uint8_t* out);

/* compute_tessellation_from_file (src/input/mod.rs:518-552): default parse
 * options and radii; selections NULL = no grouping, else at least two. */
int vltr_compute_tessellation_from_file(vltr_context* ctx, const char* path, double probe,
                                        const vltr_periodic_box* periodic_box, int with_cell_vertices,
                                        const char* const* selections, size_t n_selections, vltr_result** out);

/* The CLI's JSON document (src/main.rs:22-34, serde_json pretty printer):
 * the result's fields, per-ball sas_areas / volumes (null without a cell) and
 * the three totals. Free the text with vltr_string_free. */
int vltr_result_to_json(const vltr_result* r, char** text, size_t* len);

/* GraphicsWriter::from_balls + write_pymol (src/graphics.rs:45-178): PyMOL
 * CGO script of the balls and of every contact polygon (arcs sampled every
 * angle_step radians), non-periodic, groups optional. */
int vltr_pymol_graphics(vltr_context* ctx, const vltr_ball* balls, size_t n, double probe, const int32_t* groups,
                        size_t n_groups, double angle_step, const char* object_name, char** text, size_t* len);

void vltr_string_free(char* text);

#ifdef __cplusplus
}
#endif

#endif
