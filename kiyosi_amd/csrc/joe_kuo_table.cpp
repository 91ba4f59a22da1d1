// Embeds the Joe-Kuo new-joe-kuo-6.21201 direction-number records
// (kiyosi_amd/data/joe_kuo_6_21201.u32, little-endian uint32 words; see
// kiyosi_amd/data/extract_joe_kuo.py for provenance and licence) into the
// library as kg_joe_kuo_words .. kg_joe_kuo_end.
#ifndef KG_JOE_KUO_PATH
#error "KG_JOE_KUO_PATH must name the direction-number file"
#endif
#define KG_STR2(x) #x
#define KG_STR(x) KG_STR2(x)
__asm__(".section .rodata\n"
        ".balign 16\n"
        ".global kg_joe_kuo_words\n"
        "kg_joe_kuo_words:\n"
        ".incbin \"" KG_STR(KG_JOE_KUO_PATH) "\"\n"
        ".global kg_joe_kuo_end\n"
        "kg_joe_kuo_end:\n"
        ".previous\n");
