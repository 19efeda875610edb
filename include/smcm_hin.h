/*
 * smcm_hin.h - host-side heterogeneous information network (HIN) utilities.
 *
 * A HIN is held as a global adjacency CSR: every vertex v has a type vtype[v] and an ordered
 * neighbour list adj_nbr[adj_ptr[v] .. adj_ptr[v+1]) whose entries carry the edge type of the
 * edge they came from (adj_et).  Neighbour order is file order; it is the order in which the
 * row-wise sparsity estimator multiplies densities, so it is never changed by this library.
 *
 * Replaces (as a C ABI, no C++ types crossing it):
 *   DataReader            efficient/resmcm.cpp:49-177   (text loader: graph.txt/vertex.txt/edge.txt)
 *   HeteroGraph           efficient/resmcm.cpp:181-240  (types, dimensions, bins)
 *   read_metapathesN      efficient/resmcm.cpp:339-357  (meta-path list file)
 *   MetaPath(string)      efficient/resmcm.cpp:269-299  (meta-path parser)
 *   MetaPath::toString    efficient/resmcm.cpp:301-308
 * plus the seeded synthetic generators used for the benchmark workloads (no counterpart in the
 * reference, whose datasets are download-only).
 *
 * All functions return 0 on success and a negative errno-style code on failure; the message
 * of the last failure on the calling thread is available from hin_last_error().
 */
#ifndef SMCM_HIN_H
#define SMCM_HIN_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define HIN_MAX_TYPES 32

typedef struct hin_graph {
  int32_t n;          /* number of vertices (global ids 0..n-1) */
  int64_t m;          /* number of adjacency entries (directed edges) */
  int32_t ntypes;     /* max vertex type + 1 */
  int32_t* vtype;     /* [n] vertex type */
  int64_t* adj_ptr;   /* [n+1] */
  int32_t* adj_nbr;   /* [m] neighbour global id, file order */
  int32_t* adj_et;    /* [m] edge type of the entry */
  int32_t dims[HIN_MAX_TYPES];        /* number of vertices per type (HeteroGraph::set_dimensions) */
  int32_t type_offset[HIN_MAX_TYPES]; /* first global id of the type when ids of that type are
                                         one contiguous range, else -1 */
} hin_graph;

/* Seeded bibliographic graph (DBLP schema: paper 0, author 1, venue 2, term 3; edge types
 * P->A 0, P->V 1, P->T 2, A->P 3, V->P 4, T->P 5, as in tool/generate_DBLP.cpp:334-352).
 * Each paper gets 1+Poisson(lam_auth) distinct authors ~ Zipf(s_auth) over A, exactly one
 * venue ~ Zipf(s_venue) over V and terms_per_paper distinct terms ~ Zipf(s_term) over T
 * (nT = 0 or terms_per_paper = 0 disables terms).  Ids are type-contiguous in the order
 * P, A, V, T and neighbour lists are ascending. */
int hin_gen_bib(int32_t nA, int32_t nP, int32_t nV, int32_t nT, double lam_auth, double s_auth,
                double s_venue, int32_t terms_per_paper, double s_term, uint64_t seed,
                hin_graph** out);

/* Seeded location graph (FourSquare schema: venue 0, city 1, category 2, user 3, slot 4; edge
 * types V->Y 0, Y->V 1, V->C 2, C->V 3, V->U 4, U->V 5, V->S 6, S->V 7, as in
 * tool/generate_FourSquare.cpp:335-359).  Each user has 1+Geometric(geo_p) distinct venues
 * ~ Zipf(s_checkin); each venue has one city ~ Zipf(s_city), 1-3 distinct categories and 1-4
 * distinct slots drawn uniformly. */
int hin_gen_loc(int32_t nU, int32_t nL, int32_t nC, int32_t nY, int32_t nS, double geo_p,
                double s_checkin, double s_city, uint64_t seed, hin_graph** out);

/* Build a graph from caller arrays (copied). */
int hin_from_arrays(int32_t n, const int32_t* vtype, const int64_t* adj_ptr,
                    const int32_t* adj_nbr, const int32_t* adj_et, hin_graph** out);

/* Reference text format (readme.md:23; loader efficient/resmcm.cpp:86-176):
 *   vertex.txt  "v vt"            one line per vertex
 *   edge.txt    "e et"            one line per edge id
 *   graph.txt   "v nb e nb e ..." one line per vertex, neighbours in file order
 * The writer gives every adjacency entry its own edge id and writes a line for every vertex. */
int hin_write_text(const hin_graph* g, const char* graph_file, const char* vertex_file,
                   const char* edge_file);
int hin_read_text(const char* graph_file, const char* vertex_file, const char* edge_file,
                  hin_graph** out);

void hin_free(hin_graph* g);

/* Meta-path parsing with the reference's rules (efficient/resmcm.cpp:269-299): any non-digit
 * character separates numbers; numbers alternate vertex type, edge type, vertex type, ...
 * Writes up to cap vertex types and cap-1 edge types; *L receives the number of matrices.
 * Returns -EINVAL when the counts do not satisfy #vtypes == #etypes + 1. */
int hin_parse_metapath(const char* s, int32_t* vtypes, int32_t* etypes, int32_t cap, int32_t* L);
/* Every meta-path with L edges over a schema, as tool/generate_DBLP.cpp:286-304 (and
 * generate_FourSquare.cpp / generate_IMDB.cpp) enumerate them: depth-first from each start type in
 * starts[] order, neighbours of a type in the order of the relations (rel_src[r] --rel_et[r]-->
 * rel_dst[r]) listed for it.  Writes at most cap meta-paths (out_vts: cap*(L+1), out_ets: cap*L);
 * *count receives the total. */
int hin_enumerate_metapaths(const int32_t* rel_src, const int32_t* rel_et, const int32_t* rel_dst,
                            int32_t nrel, const int32_t* starts, int32_t nstarts, int32_t L,
                            int32_t* out_vts, int32_t* out_ets, int64_t cap, int64_t* count);
/* "vt-et-vt-...-vt" as MetaPath::toString; returns the length written (excluding NUL). */
int hin_metapath_to_string(const int32_t* vtypes, const int32_t* etypes, int32_t L, char* buf,
                           int32_t cap);

const char* hin_last_error(void);

#ifdef __cplusplus
}
#endif
#endif
