// rader_gen_impl.h -- shape table and launches of one generated list of k_rowrader
// shapes (rader_shapes.h).  Included once by each rowpass_rader_gen_*.hip with
// SPFFT_GEN_LIST (the list) and SPFFT_GEN_FN(name) (the name with the file's suffix)
// defined, inside namespace spfft.

bool SPFFT_GEN_FN(rader_gen_shape)(int i, RaderShape* out) {
    static const RaderShape kShapes[] = {
#define SPFFT_SHAPE(m1, nt, r1, r2, r3, ns) {m1, nt, {r1, r2, r3}, rader_lds_bytes_of(m1, nt, r1, r2, r3, ns), 0, ns},
        SPFFT_GEN_LIST(SPFFT_SHAPE)
#undef SPFFT_SHAPE
    };
    if (i < 0 || i >= (int)(sizeof(kShapes) / sizeof(kShapes[0]))) return false;
    *out = kShapes[i];
    return true;
}

hipError_t SPFFT_GEN_FN(launch_rader_gen)(const RowPassArgs& a, hipStream_t s, bool prepare_only, bool* found) {
    *found = true;
#define SPFFT_CASE(m1, nt, r1, r2, r3, ns) \
    if (rader_shape_is(a, m1, nt, r1, r2, r3, ns)) return launch_rader_fp32<m1, nt, r1, r2, r3, ns>(a, s, prepare_only);
    SPFFT_GEN_LIST(SPFFT_CASE)
#undef SPFFT_CASE
    *found = false;
    return hipErrorInvalidValue;
}

hipError_t SPFFT_GEN_FN(launch_rader_gen_batch)(const RowPassArgs& a, const RowPassArgs* descs, const int2* work,
                                                int n, hipStream_t s, bool* found) {
    *found = true;
#define SPFFT_CASE(m1, nt, r1, r2, r3, ns)         \
    if (rader_shape_is(a, m1, nt, r1, r2, r3, ns)) \
        return launch_rader_fp32_batch<m1, nt, r1, r2, r3, ns>(descs, work, n, s);
    SPFFT_GEN_LIST(SPFFT_CASE)
#undef SPFFT_CASE
    *found = false;
    return hipErrorInvalidValue;
}
