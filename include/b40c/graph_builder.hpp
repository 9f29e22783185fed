/*
 * C++ facade over the graph-file readers of the b40c_mi355x C ABI, with the
 * reference's builder names (b40c/graph/builder/{market,metis,dimacs}.cuh):
 * BuildMarketGraph (market.cuh:172-205), BuildMetisGraph (metis.cuh:184-212),
 * BuildDimacsGraph (dimacs.cuh:230-283) and WriteMetisGraph (metis.cuh:252-283).
 * Where the reference fills a host CsrGraph and returns 0 / -1, these build a
 * device b40c_graph (to hand to CsrProblem::FromGraph) and return a
 * b40c_status; a NULL filename reads stdin (writes stdout), as there.
 */
#pragma once

#include "b40c_mi355x.h"

namespace b40c {
namespace graph {
namespace builder {

inline b40c_status BuildMarketGraph(const char* market_filename, b40c_graph** out, bool undirected) {
    return b40c_graph_read(market_filename, B40C_FILE_MARKET, undirected ? 1 : 0, 0, out);
}

inline b40c_status BuildMetisGraph(const char* metis_filename, b40c_graph** out) {
    return b40c_graph_read(metis_filename, B40C_FILE_METIS, 0, 0, out);
}

// splice > 0 reads dimacs_filename.0 .. .<splice-1>, merged by row.
inline b40c_status BuildDimacsGraph(const char* dimacs_filename, b40c_graph** out, bool undirected, int splice) {
    return b40c_graph_read(dimacs_filename, B40C_FILE_DIMACS, undirected ? 1 : 0, splice, out);
}

inline b40c_status WriteMetisGraph(const char* metis_filename, const b40c_graph* g) {
    return b40c_graph_write_metis(g, metis_filename);
}

}  // namespace builder
}  // namespace graph
}  // namespace b40c
