# The MI355X model for the reference's model registry (CMakeLists.txt:302-313,
# cmake/register_models.cmake:115-146).  The model sources are plain C++17 over
# the clover_mi355x C ABI; the gfx950 kernels live in libclover_mi355x.so, built
# by cloverleaf_amd/csrc/Makefile.
register_flag_required(CLOVER_MI355X_ROOT
        "Path to the clover_mi355x repository (include/ and cloverleaf_amd/libclover_mi355x.so)")

macro(setup)
    set(CMAKE_CXX_STANDARD 17)
    include_directories(${CLOVER_MI355X_ROOT}/include)
endmacro()

macro(setup_target NAME)
    target_link_libraries(${NAME} PUBLIC ${CLOVER_MI355X_ROOT}/cloverleaf_amd/libclover_mi355x.so)
    set_target_properties(${NAME} PROPERTIES BUILD_RPATH ${CLOVER_MI355X_ROOT}/cloverleaf_amd)
endmacro()
