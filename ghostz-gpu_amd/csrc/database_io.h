// database_io.h -- the on-disk database of `ghostz-gpu db`: <prefix>.inf and,
// per chunk i, <prefix>_<i>.{inf,off,nam,seq,scp,sdp} (src/database.h:150-152,
// src/database_chunk.h:103-123), read into and written from the in-memory
// Database the search uses.
//
// Layout (all integers little-endian, written field by field without
// padding; "vector" = uint64 count followed by the elements):
//   <prefix>.inf    int32 chunks, uint32 max sequence length, uint64 total
//                   residues, uint64 sequences, uint8 delimiter
//   _i.inf          uint32 sequences, uint32 concatenated length
//   _i.off          uint32 x (sequences + 1): subject starts, then the end
//   _i.nam          one name per line
//   _i.seq          uint8 codes: delimiter, subject, delimiter, ...
//   _i.scp          uint8 delimiter, uint32 L, vector<uint8> reduced map,
//                   hash function
//   _i.sdp          uint8 clustering, uint8 delimiter, uint32 L,
//                   hash function + dense index (hash -> positions),
//                   hash function + dense index (hash -> (cluster, position))
//                   + perfect index (cluster -> member positions),
//                   float cluster rate, vector<uint8> reduced map
// hash function = uint8 usable, uint8 max code + 1, int32 score threshold,
// uint32 max length, uint32 shift, uint32 max hash, vector<uint8> code map
// (+1), vector<int32> code scores; dense index = uint64 n, n x (uint32
// hash, uint32 slot), vector<uint32> slot offsets, vector<value>; perfect
// index = uint32 max key, vector<uint32> offsets, vector<uint32> values.
#pragma once

#include <string>

#include "database.h"

namespace gz {

// Database::Load (src/database.h:229-237) with every chunk resident
// (DatabaseChunk::Load, src/database_chunk.h:303-400).  False with *err set
// when a file is missing, truncated or inconsistent.
bool LoadDatabase(const std::string &prefix, Database *db, std::string *err);

// Database::Save (src/database.h:240-251): the same file set.
bool SaveDatabase(const Database &db, const std::string &prefix, std::string *err);

}  // namespace gz
