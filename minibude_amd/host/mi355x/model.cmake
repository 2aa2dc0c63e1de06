# model.cmake of the mi355x model (src/mi355x/model.cmake in a miniBUDE
# tree).  The model is host C++ only: any C++17 compiler builds it; the GPU
# code lives in the prebuilt libbude_mi355x.so it links.
register_flag_required(MI355X_ROOT
        "Absolute path to the bude-mi355x tree (include/bude_mi355x.h, minibude_amd/lib/libbude_mi355x.so)")

macro(setup)
    include_directories(${MI355X_ROOT}/include)
    register_link_library(${MI355X_ROOT}/minibude_amd/lib/libbude_mi355x.so)
endmacro()
