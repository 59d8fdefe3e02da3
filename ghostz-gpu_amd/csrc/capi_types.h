// capi_types.h -- the objects behind the opaque query and database handles
// of the C ABI (shared by the C ABI's source files).
#pragma once

#include <vector>

#include "database.h"
#include "queries.h"

struct gz_queries_s {
  gz::QuerySet q;
  std::vector<const char *> name_ptrs;
};

struct gz_database_s {
  gz::Database db;
  std::vector<std::vector<const char *> > name_ptrs;
};
