// The perm-path hook the harness expects; the MI355X backend computes its
// own ordering, so the path is ignored (TEST INFRASTRUCTURE ONLY).
const char* g_ref_perm_path = nullptr;
