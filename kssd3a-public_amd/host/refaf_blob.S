/* Embeds refaf_hgb.bin (the calibrated-ANI model tables) in the host library. */
	.section .rodata
	.balign 8
	.globl kssd_refaf_blob
	.globl kssd_refaf_blob_end
kssd_refaf_blob:
	.incbin "refaf_hgb.bin"
kssd_refaf_blob_end:
	.section .note.GNU-stack,"",@progbits
